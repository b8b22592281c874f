"""Edge cases of the trace step on the GPU against the oracle: a scene with no
geometry (only the environment, or nothing), ragged extents, degenerate
triangles, a frame split into several chunks of paths, and ray queries with
no rays, empty intervals and rays that start beyond the scene.
"""

import numpy as np
import pytest

import oracle as O
import vkrt_amd as V
from vkrt_amd import scenes

pytestmark = pytest.mark.gpu


def render_pair(scene, w, h, spp, depth=6, mode=V.RENDER_MODE_RGB, rows=None, bvh=False):
    r = V.Renderer(w, h)
    scenes.load(r, scene)
    r.configure(spp=spp, max_depth=depth, min_depth=4, mode=mode, spectral=V.SPECTRAL_HERO, rng=V.RNG_PCG32, seed=7)
    r.start(w, h, 0)
    orc = O.OracleScene(r.export(), use_bvh=bvh)
    r.draw()
    gpu = r.accumulation()
    st = r.stats()
    ref, bounces = orc.render(w, h, precision=32, rows=rows)
    orc.close()
    r.close()
    return gpu, ref, st, bounces


def empty_scene(env=True):
    s = scenes.Scene()
    s.camera = ((0.0, -3.0, 0.5), (0.0, 0.0, 0.0), (0.0, 0.0, 1.0), 50.0)
    if env:
        s.environment = scenes.sky_environment(64, 32, sun_radiance=20.0, sun_diameter_deg=8.0)
    return s


@pytest.mark.parametrize("mode", [V.RENDER_MODE_RGB, V.RENDER_MODE_SPECTRAL])
def test_scene_without_geometry_shows_the_environment(mode):
    """No triangles: every camera ray misses and gathers the environment."""
    gpu, ref, st, bounces = render_pair(empty_scene(), 37, 23, 3, mode=mode)
    err = O.relmse(gpu[..., :3], ref[..., :3])
    print(f"relMSE {err:.3e}, bounces gpu {st.extensionRays} oracle {bounces}")
    assert st.triangleCount == 0
    assert st.extensionRays == bounces == 37 * 23 * 3
    assert np.all(gpu[..., 3] == 3) and gpu[..., :3].mean() > 0
    assert err < 1e-12


def test_scene_without_geometry_or_environment_is_black():
    s = empty_scene(env=False)
    s.environment_color, s.environment_strength = (0.0, 0.0, 0.0), 0.0
    gpu, ref, _, _ = render_pair(s, 16, 9, 2)
    assert np.all(gpu[..., :3] == 0.0) and np.all(ref[..., :3] == 0.0)
    assert np.all(gpu[..., 3] == 2)


def test_ragged_extent_and_odd_sample_count():
    """A 61 x 37 frame at 5 spp: nothing in the pipeline assumes sizes that are
    multiples of a block, a wavefront or a chunk."""
    gpu, ref, st, bounces = render_pair(scenes.cornell_box(), 61, 37, 5, depth=8)
    err = O.relmse(gpu[..., :3], ref[..., :3])
    print(f"relMSE {err:.3e}, bounces gpu {st.extensionRays} oracle {bounces}")
    assert np.all(gpu[..., 3] == 5)
    assert err < 1e-12 and st.extensionRays == bounces


def test_degenerate_triangles_change_nothing():
    """Zero-area triangles (a repeated vertex, three collinear vertices) added
    to the Cornell box are never hit: the image matches the oracle and the
    image without them (the BVH differs, so a ray on an edge shared by two
    coplanar triangles may report the other one: equal to rounding)."""
    base = scenes.cornell_box()
    s = scenes.cornell_box()
    pos = np.array([[0.5, 0.5, 0.5], [0.5, 0.5, 0.5], [0.6, 0.4, 0.5],
                    [0.2, 0.2, 0.2], [0.4, 0.4, 0.4], [0.6, 0.6, 0.6]], np.float32)
    nrm = np.tile(np.array([0, 0, 1], np.float32), (6, 1))
    s.meshes.append(scenes.MeshData(V.make_vertices(pos, nrm), np.array([[0, 1, 2], [3, 4, 5]], np.uint32), 0))
    a, ref_a, _, _ = render_pair(base, 48, 48, 4, depth=8)
    b, ref_b, _, _ = render_pair(s, 48, 48, 4, depth=8)
    assert O.relmse(b[..., :3], ref_b[..., :3]) < 1e-12
    assert O.relmse(b[..., :3], a[..., :3]) < 1e-12 and np.array_equal(a[..., 3], b[..., 3])


def test_frame_split_into_two_chunks_matches_the_oracle():
    """1024 x 1024 at 160 spp is 167.8M paths: more than one chunk of 2^27
    (128 whole samples of every pixel per chunk), so the frame runs as two
    chunks whose per-pixel sums must join in sample order. A band of rows is
    compared with the oracle, which renders every sample of those rows."""
    w = h = 1024
    rows = (500, 503)
    gpu, ref, st, _ = render_pair(scenes.cornell_box(), w, h, 160, depth=4, rows=rows)
    g, o = gpu[rows[0]:rows[1], :, :3], ref[rows[0]:rows[1], :, :3]
    err = O.relmse(g, o)
    print(f"rows {rows}: relMSE {err:.3e}; GPU bounces (whole frame) {st.extensionRays}, paths {st.paths}")
    assert st.paths == w * h * 160
    assert np.all(gpu[..., 3] == 160)
    assert err < 1e-12


def test_ray_queries_edge_cases():
    """No rays, an empty interval (tmin > tmax), rays that start beyond the
    scene and point away, and a ray through the box's centre."""
    r = V.Renderer(16, 16)
    scenes.load(r, scenes.cornell_box())
    r.configure(spp=1, max_depth=4, min_depth=4, mode=V.RENDER_MODE_RGB, rng=V.RNG_PCG32, seed=1)
    r.start(16, 16, 0)
    m, p, t, b = r.trace_rays(np.zeros((0, 8), np.float32))
    assert len(m) == 0
    rays = np.array([
        [0.5, 0.5, 0.5, 2.0, 0.0, 0.0, 1.0, 1.0],      # tmin > tmax
        [0.5, -5.0, 0.5, 0.0, 0.0, -1.0, 0.0, 1e30],   # outside, pointing away
        [50.0, 50.0, 50.0, 0.0, 1.0, 0.0, 0.0, 1e30],  # far outside, parallel to an axis
        [0.5, 0.5, 0.5, 0.0, 0.0, 0.0, 1.0, 1e30],     # from the centre up: the light or the ceiling
    ], np.float32)
    m, p, t, b = r.trace_rays(rays)
    om, op, ot, _ = O.OracleScene(r.export(), use_bvh=False).trace_rays(rays)
    assert list(m[:3]) == [V.VKRT_INVALID_INDEX] * 3
    assert m[3] != V.VKRT_INVALID_INDEX
    assert np.array_equal(m, om) and np.array_equal(p, op)
    assert t[3] == pytest.approx(float(ot[3]), rel=1e-6)
    vis = r.trace_shadow_rays(rays[:3])
    assert np.all(vis == 1)  # nothing occludes an empty interval or a ray that leaves the scene
    r.close()
