"""VKRT_traceRaysDeviceNoAlloc, the allocation-free ray query, against
VKRT_traceRays on the same scene and rays: mesh and primitive indices, hit
distances and barycentrics bit for bit, on two calls in a row with the second
smaller (it writes its own count of entries and leaves the rest of the
buffers as the first call left them), with the optional outputs left out,
and its refusals (more rays than its maximum, a scene changed since its last
upload). Both queries run the same traversal on the same keys, so equality is
exact; the host path's own agreement with the oracle is covered with the
other ray-query tests.
"""

import numpy as np
import pytest

import vkrt_amd as V
from vkrt_amd import scenes

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

SENTINEL = -1515870811  # 0xA5A5A5A5 as int32


def _renderer(scene, w=64, h=64):
    r = V.Renderer(w, h)
    scenes.load(r, scene)
    r.configure(spp=1, max_depth=8, min_depth=4, mode=V.RENDER_MODE_RGB, rng=V.RNG_PCG32, seed=1)
    r.start(w, h, 0)
    return r


def _rays_in_box(rng, n, lo, hi, tmin=1e-3):
    rays = np.zeros((n, 8), np.float32)
    rays[:, 0:3] = rng.uniform(lo, hi, (n, 3))
    d = rng.normal(size=(n, 3))
    rays[:, 4:7] = d / np.linalg.norm(d, axis=1, keepdims=True)
    rays[:, 3], rays[:, 7] = tmin, 1e4
    return rays


def _rays_at_sphere(rng, n):
    """Rays from a shell of radius 3 aimed at points inside the unit sphere."""
    o = rng.normal(size=(n, 3))
    o *= 3.0 / np.linalg.norm(o, axis=1, keepdims=True)
    d = rng.uniform(-0.8, 0.8, (n, 3)) - o
    rays = np.zeros((n, 8), np.float32)
    rays[:, 0:3] = o
    rays[:, 4:7] = d / np.linalg.norm(d, axis=1, keepdims=True)
    rays[:, 3], rays[:, 7] = 0.0, 1e4
    return rays


class DeviceOutputs:
    """Output buffers on cuda:0 with 16 sentinel words past the end."""

    def __init__(self, n):
        dev = torch.device("cuda:0")
        self.n = n
        self.mesh = torch.full((n + 16,), SENTINEL, dtype=torch.int32, device=dev)
        self.prim = torch.full((n + 16,), SENTINEL, dtype=torch.int32, device=dev)
        self.t = torch.full((n + 16,), SENTINEL, dtype=torch.int32, device=dev)
        self.bary = torch.full((2 * n + 16,), SENTINEL, dtype=torch.int32, device=dev)

    def host(self):
        torch.cuda.synchronize()
        return [x.cpu().numpy().view(np.uint32) for x in (self.mesh, self.prim, self.t, self.bary)]


def _noalloc(r, rays, out, with_t=True):
    dr = torch.from_numpy(np.ascontiguousarray(rays)).to("cuda:0")
    torch.cuda.synchronize()
    r.trace_rays_device_noalloc(dr.data_ptr(), len(rays), out.mesh.data_ptr(), out.prim.data_ptr(),
                                out.t.data_ptr() if with_t else 0, out.bary.data_ptr() if with_t else 0)
    torch.cuda.synchronize()


def _host(r, rays):
    m, p, t, b = r.trace_rays(rays)
    return m, p, t.view(np.uint32), b.reshape(-1).view(np.uint32)


def _check_prefix(got, want, n):
    m, p, t, b = got
    hm, hp, ht, hb = want
    assert np.array_equal(m[:n], hm), "mesh"
    assert np.array_equal(p[:n], hp), "primitive"
    assert np.array_equal(t[:n], ht), "t"
    assert np.array_equal(b[:2 * n], hb), "barycentrics"


def _two_calls(r, first, second):
    n1, n2 = len(first), len(second)
    assert n2 < n1
    want1 = _host(r, first)  # also uploads the scene, which the allocation-free query requires
    out = DeviceOutputs(n1)
    _noalloc(r, first, out)
    got1 = out.host()
    _check_prefix(got1, want1, n1)
    for a in got1:
        assert np.all(a[-16:] == np.uint32(0xA5A5A5A5)), "wrote past the end"
    hits = np.mean(want1[0] != V.VKRT_INVALID_INDEX)
    # second call, smaller, into the same buffers
    _noalloc(r, second, out)
    got2 = out.host()
    want2 = _host(r, second)
    _check_prefix(got2, want2, n2)
    m1, p1, t1, b1 = got1
    m2, p2, t2, b2 = got2
    assert np.array_equal(m2[n2:], m1[n2:]) and np.array_equal(p2[n2:], p1[n2:])
    assert np.array_equal(t2[n2:], t1[n2:]) and np.array_equal(b2[2 * n2:], b1[2 * n2:])
    return hits


def test_noalloc_matches_trace_rays_on_alpha_tested_cornell_box():
    rng = np.random.default_rng(21)
    r = _renderer(scenes.alpha_cornell())
    try:
        hits = _two_calls(r, _rays_in_box(rng, 131072, 0.02, 0.98), _rays_in_box(rng, 3001, 0.02, 0.98))
        print(f"alpha Cornell: {hits:.4f} of the first call's rays hit", flush=True)
        assert hits > 0.5
    finally:
        r.close()


def test_noalloc_matches_trace_rays_on_displaced_sphere():
    rng = np.random.default_rng(22)
    r = _renderer(scenes.displaced_sphere(229))
    try:
        hits = _two_calls(r, _rays_at_sphere(rng, 1 << 20), _rays_at_sphere(rng, 65537))
        print(f"displaced sphere: {hits:.4f} of the first call's rays hit", flush=True)
        assert hits > 0.5
    finally:
        r.close()


def test_noalloc_without_optional_outputs_and_refusals():
    rng = np.random.default_rng(23)
    r = _renderer(scenes.cornell_box())
    try:
        rays = _rays_in_box(rng, 4096, 0.05, 0.95)
        want = _host(r, rays)
        out = DeviceOutputs(len(rays))
        _noalloc(r, rays, out, with_t=False)
        m, p, t, b = out.host()
        assert np.array_equal(m[:4096], want[0]) and np.array_equal(p[:4096], want[1])
        assert np.all(t == np.uint32(0xA5A5A5A5)) and np.all(b == np.uint32(0xA5A5A5A5))
        # beyond the per-call maximum: refused on the host before any launch
        with pytest.raises(V.VkrtError):
            r.trace_rays_device_noalloc(out.mesh.data_ptr(), (1 << 30) + 1, out.mesh.data_ptr(),
                                        out.prim.data_ptr())
        # a scene changed since its last upload: refused (uploading allocates)
        r.set_material(0, r.material(0))
        with pytest.raises(V.VkrtError):
            _noalloc(r, rays, out)
        want = _host(r, rays)  # uploads the scene
        _noalloc(r, rays, out)
        _check_prefix(out.host(), want, len(rays))
    finally:
        r.close()
