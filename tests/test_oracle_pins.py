"""Pins the CPU oracle before it is trusted as the parity checker.

The reference ships no golden vectors for this path, so each oracle function
is checked against something outside itself: a published known-answer vector
(PCG32, CIE 1931 table), a textbook routine (Fresnel equations, Bradford
adaptation + IEC 61966-2-1 sRGB matrix, Abbe number), a closed form (camera
frustum, barycentrics, cosine pdf, furnace and closed-sphere radiance), an
invariant (alias tables, pdf normalisation, sample/eval agreement,
reciprocity, watertightness) or brute force (BVH against all triangles).
"""

import ctypes as C
import json
import math
import os

import numpy as np
import pytest

import oracle as O
import vkrt_amd as V
from vkrt_amd import scenes

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def golden(name):
    with open(os.path.join(GOLDEN, name)) as f:
        return json.load(f)


def cpu_renderer(width=32, height=32):
    r = V.Renderer(init_device=False)
    r.start(width, height, 0)
    return r


def faceted_sphere(freq, radius=1.0, inward=False):
    """Sphere with one vertex triple per face and flat normals (no shading-normal tilt)."""
    verts, tris = scenes.geodesic_sphere(freq)
    p = (verts[tris] * radius).reshape(-1, 3)
    a, b, c = p[0::3], p[1::3], p[2::3]
    n = np.cross(b - a, c - a)
    n /= np.linalg.norm(n, axis=1, keepdims=True)
    outward = np.sign(np.sum(n * (a + b + c), axis=1))[:, None]
    n *= outward
    idx = np.arange(len(p), dtype=np.uint32).reshape(-1, 3)
    flip = (outward[:, 0] < 0) ^ inward
    idx[flip] = idx[flip][:, [0, 2, 1]]
    if inward:
        n = -n
    return V.make_vertices(p.astype(np.float32), np.repeat(n, 3, axis=0).astype(np.float32)), idx


# ------------------------------------------------------------------ RNG --
def test_pcg32_published_known_answers():
    kat = golden("pcg32_kat.json")
    got = O.pcg32_raw(kat["init_state"], kat["init_seq"], len(kat["outputs"]))
    assert [int(x, 16) for x in kat["outputs"]] == [int(v) for v in got]


@pytest.mark.parametrize("policy", [V.RNG_HASH, V.RNG_PCG32])
def test_uniforms_are_24_bit_and_uniform(policy):
    u = np.concatenate([O.rand_sequence(policy, 1, px, 7, 0, 0, 4096, precision=64) for px in range(16)])
    assert u.min() >= 0.0 and u.max() < 1.0
    assert np.all(np.floor(u * 2 ** 24) == u * 2 ** 24)  # random.slang:16-19 keeps 24 bits
    counts, _ = np.histogram(u, bins=64, range=(0, 1))
    expected = len(u) / 64
    chi2 = float(np.sum((counts - expected) ** 2 / expected))
    assert chi2 < 120.0  # 63 dof; p ~ 1e-5
    # lag-1 correlation
    assert abs(np.corrcoef(u[:-1], u[1:])[0, 1]) < 0.02


M32 = 0xFFFFFFFF


def lowbias32_inverse(x):
    """Wellons' published inverse of lowbias32 (nullprogram.com/blog/2018/07/31,
    'lowbias32_r'): multipliers 0x43021123 and 0x1d69e2a5 undo 0x846ca68b and
    0x7feb352d, and x ^= x>>15 ^ x>>30 undoes x ^= x>>15."""
    x ^= x >> 16
    x = (x * 0x43021123) & M32
    x ^= (x >> 15) ^ (x >> 30)
    x = (x * 0x1D69E2A5) & M32
    x ^= x >> 16
    return x


def test_lowbias32_inverse_constants_are_the_modular_inverses():
    assert pow(0x7FEB352D, -1, 1 << 32) == 0x1D69E2A5
    assert pow(0x846CA68B, -1, 1 << 32) == 0x43021123


def test_hash_is_lowbias32():
    """random.slang:7-14. Inverting the oracle's hash with the published inverse
    must give back every input; a wrong multiplier, shift or order breaks it."""
    rng = np.random.default_rng(11)
    xs = [0, 1, 2, 0x9E3779B9, M32, 0x80000000] + [int(v) for v in rng.integers(0, 1 << 32, 2000, dtype=np.uint64)]
    for x in xs:
        assert lowbias32_inverse(int(O.oracle_hash(x))) == x
    # the hash is not the identity or an xorshift alone
    assert O.oracle_hash(1) != 1 and O.oracle_hash(0) == 0


@pytest.mark.parametrize("px,py,frame,sample", [(0, 0, 0, 0), (3, 4, 0, 0), (1919, 1079, 7, 63),
                                                (255, 0, 123456, 15), (17, 999, 0xFFFFFFFF, 0xFFFFFFFF)])
def test_hash_pixel_seed_and_steps(px, py, frame, sample):
    """initPixelSeed (random.slang:21-27): inverse(seed) is the xor of the four
    products. rand (random.slang:16-19): inverse(next) = state + 0x9e3779b9 and
    the uniform is the low 24 bits of the new state over 2^24."""
    n = 9
    states = [int(s) for s in O.hash_states(px, py, frame, sample, n)]
    want = ((px * 73856093) ^ (py * 19349663) ^ (frame * 83492791) ^ (sample * 2654435761)) & M32
    assert lowbias32_inverse(states[0]) == want
    for k in range(1, n):
        assert lowbias32_inverse(states[k]) == (states[k - 1] + 0x9E3779B9) & M32
    u = O.rand_sequence(V.RNG_HASH, 1, px, py, frame, sample, n - 1, precision=64)
    assert np.array_equal(u, np.array([(s & 0xFFFFFF) / 16777216.0 for s in states[1:]]))


def test_streams_differ_by_pixel_frame_and_sample():
    base = O.rand_sequence(V.RNG_PCG32, 1, 3, 4, 0, 0, 8)
    for args in ((4, 4, 0, 0), (3, 5, 0, 0), (3, 4, 1, 0), (3, 4, 0, 1)):
        assert not np.array_equal(base, O.rand_sequence(V.RNG_PCG32, 1, *args, 8))
    assert not np.array_equal(base, O.rand_sequence(V.RNG_PCG32, 2, 3, 4, 0, 0, 8))


# --------------------------------------------------------------- camera --
def test_camera_rays_follow_the_perspective_frustum():
    r = cpu_renderer(64, 48)
    r.set_camera((0.0, -2.0, 0.0), (0.0, 0.0, 0.0), (0.0, 0.0, 1.0), 40.0)
    b = r.export()
    sd = b.sceneData
    out = np.zeros(6)
    fn = O.lib().oracle_camera_ray_f64

    def ray(px, py, jx, jy):
        fn(C.cast(sd, C.c_void_p), px, py, jx, jy, O.ptr(out))
        d = out[3:].copy()
        return out[:3].copy(), d / np.linalg.norm(d)

    o, d = ray(32, 24, -0.5, -0.5)  # exact image centre
    assert np.allclose(o, (0, -2, 0), atol=1e-12)
    assert np.allclose(d, (0, 1, 0), atol=1e-9)
    t = math.tan(math.radians(20.0))
    aspect = 64 / 48
    _, d = ray(0, 0, -0.5, -0.5)  # top-left image corner: left (-x) and up (+z)
    want = np.array([-aspect * t, 1.0, t])
    assert np.allclose(d, want / np.linalg.norm(want), atol=1e-9)
    _, d = ray(63, 47, 0.5, 0.5)  # bottom-right corner
    want = np.array([aspect * t, 1.0, -t])
    assert np.allclose(d, want / np.linalg.norm(want), atol=1e-9)
    r.close()


# --------------------------------------------------------- intersection --
TRI = np.array([0, 0, 0, 1, 0, 0, 0, 1, 0], np.float32)


@pytest.mark.parametrize("precision", [32, 64])
def test_intersection_closed_form(precision):
    rng = np.random.default_rng(3)
    for _ in range(200):
        x, y = rng.uniform(0.01, 0.49, 2)
        h = rng.uniform(0.5, 3.0)
        for sign in (1.0, -1.0):  # both faces hit: no culling
            hit, tuv = O.intersect(TRI, [x, y, sign * h, 0.0, 0.0, 0.0, -sign, 100.0], precision)
            assert hit
            assert tuv[0] == pytest.approx(h, rel=1e-6)
            assert tuv[1] == pytest.approx(x, abs=1e-6)  # weight of v1
            assert tuv[2] == pytest.approx(y, abs=1e-6)  # weight of v2
    assert not O.intersect(TRI, [0.6, 0.6, 1, 0, 0, 0, -1, 100], precision)[0]
    assert not O.intersect(TRI, [0.2, 0.2, 1, 0, 0, 0, -1, 0.5], precision)[0]  # beyond tmax
    assert not O.intersect(TRI, [0.2, 0.2, 1, 0, 0, 0, 1, 100], precision)[0]   # pointing away


def test_intersection_is_watertight_on_shared_edges_and_vertices():
    """Rays aimed at the diagonal shared by two triangles of a quad, and at the
    centre vertex of a 6-triangle fan, never pass between them."""
    quad = [np.array([0, 0, 0, 1, 0, 0, 1, 1, 0], np.float32), np.array([0, 0, 0, 1, 1, 0, 0, 1, 0], np.float32)]
    rng = np.random.default_rng(5)
    for _ in range(3000):
        s = np.float32(rng.uniform(0.05, 0.95))
        target = np.array([s, s, 0.0])
        d = rng.normal(size=3)
        d[2] = -abs(d[2]) - 0.05
        d /= np.linalg.norm(d)
        o = (target - d * rng.uniform(0.5, 5.0)).astype(np.float32)
        ray = [o[0], o[1], o[2], 0.0, d[0], d[1], d[2], 100.0]
        assert any(O.intersect(t, ray, 32)[0] for t in quad)
    ang = np.linspace(0, 2 * np.pi, 7)[:-1]
    ring = np.stack([np.cos(ang), np.sin(ang), np.zeros(6)], 1)
    fan = [np.concatenate([[0, 0, 0], ring[i], ring[(i + 1) % 6]]).astype(np.float32) for i in range(6)]
    for _ in range(2000):
        d = rng.normal(size=3)
        d[2] = -abs(d[2]) - 0.05
        d /= np.linalg.norm(d)
        o = (-d * rng.uniform(0.5, 5.0)).astype(np.float32)
        ray = [o[0], o[1], o[2], 0.0, d[0], d[1], d[2], 100.0]
        assert any(O.intersect(t, ray, 32)[0] for t in fan)


def test_oracle_bvh_matches_brute_force():
    r = cpu_renderer()
    scenes.load(r, scenes.displaced_sphere(12))
    b = r.export()
    bvh = O.OracleScene(b, use_bvh=True)
    brute = O.OracleScene(r.export(), use_bvh=False)
    rng = np.random.default_rng(9)
    n = 4000
    o = rng.normal(size=(n, 3))
    o = o / np.linalg.norm(o, axis=1, keepdims=True) * 3.0
    d = rng.normal(size=(n, 3)) * 0.4 - o
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    rays = np.concatenate([o, np.full((n, 1), 1e-3), d, np.full((n, 1), 1e4)], 1).astype(np.float32)
    m1, p1, t1, _ = bvh.trace_rays(rays)
    m2, p2, t2, _ = brute.trace_rays(rays)
    assert (m1 != V.VKRT_INVALID_INDEX).sum() > n // 2
    assert np.array_equal(m1, m2) and np.array_equal(p1, p2) and np.array_equal(t1, t2)
    bvh.close()
    brute.close()
    r.close()


# ---------------------------------------------------------------- alias --
def alias_pmf(q, idx):
    n = len(q)
    p = q.astype(np.float64).copy()
    np.add.at(p, idx, 1.0 - q.astype(np.float64))
    return p / n


@pytest.mark.parametrize("pmf", [
    np.array([1.0]),
    np.array([0.5, 0.5]),
    np.array([0.0, 0.0, 1.0, 0.0]),
    np.random.default_rng(1).random(97),
    np.random.default_rng(2).random(1000) ** 8,
])
def test_oracle_alias_table_invariant(pmf):
    pmf = pmf / pmf.sum()
    q, idx = O.build_alias(pmf)
    assert np.all(idx < len(pmf))
    assert np.allclose(alias_pmf(q, idx), pmf, atol=2e-6)


def test_host_light_tables_reproduce_area_and_power():
    """Host alias tables (src/core/scene/lighting.c:108-433): triangles within
    a mesh by area, meshes by area x luminance."""
    r = cpu_renderer()
    rng = np.random.default_rng(4)
    tris = []
    for k in range(7):  # mesh A: 7 triangles of different areas
        s = rng.uniform(0.1, 1.0)
        tris.append(np.array([[k, 0, 0], [k + s, 0, 0], [k, s, 0]], np.float32))
    pa = np.concatenate(tris)
    va = V.make_vertices(pa, np.tile([[0, 0, 1]], (len(pa), 1)).astype(np.float32))
    ia = np.arange(len(pa), dtype=np.uint32).reshape(-1, 3)
    pb = np.array([[0, 0, 5], [1, 0, 5], [0, 1, 5]], np.float32)  # mesh B: one triangle, area 0.5
    vb = V.make_vertices(pb, np.tile([[0, 0, 1]], (3, 1)).astype(np.float32))
    ma = r.add_material(V.default_material(emissionLuminance=1.0))
    mb = r.add_material(V.default_material(emissionLuminance=3.0))
    r.set_mesh_material(r.add_mesh(va, ia), ma)
    r.set_mesh_material(r.add_mesh(vb, np.array([[0, 1, 2]], np.uint32)), mb)
    b = r.export()
    sd = b.sceneData.contents
    assert sd.emissiveMeshCount == 2 and sd.emissiveTriangleCount == 8
    meshes = [b.emissiveMeshes[i] for i in range(2)]
    areas = np.array([0.5 * np.linalg.norm(np.cross(t[1] - t[0], t[2] - t[0])) for t in tris])
    q = np.ctypeslib.as_array(b.triAliasQ, shape=(8,))
    idx = np.ctypeslib.as_array(b.triAliasIdx, shape=(8,))
    a0 = meshes[0].triOffset
    assert np.allclose(alias_pmf(q[a0:a0 + 7], idx[a0:a0 + 7]), areas / areas.sum(), atol=1e-6)
    power = np.array([areas.sum() * 1.0, 0.5 * 3.0])
    mq = np.ctypeslib.as_array(b.meshAliasQ, shape=(2,))
    midx = np.ctypeslib.as_array(b.meshAliasIdx, shape=(2,))
    assert np.allclose(alias_pmf(mq, midx), power / power.sum(), atol=1e-6)
    assert np.allclose([m.pmfMesh for m in meshes], power / power.sum(), atol=1e-6)
    r.close()


# -------------------------------------------------------------- colour --
def test_cie_fits_follow_the_published_cie1931_table():
    g = golden("cie1931_2deg.json")
    for lam, x, y, z in g["rows"]:
        got = O.cie_xyz(lam)
        assert np.allclose(got, (x, y, z), atol=0.035), (lam, got)


def test_cie_y_integral_is_the_spectral_normaliser():
    lam = np.arange(360.0, 831.0)
    xyz = np.array([O.cie_xyz(float(l)) for l in lam])
    w = np.ones_like(lam)
    w[0] = w[-1] = 0.5
    integral = (xyz * w[:, None]).sum(0)
    # the integrator divides by 106.9461715 (spectral.slang:99-113): the fitted ybar's integral
    assert integral[1] == pytest.approx(106.9461715, rel=2e-3)
    # equal-energy white -> sRGB white
    assert np.allclose(O.xyz_to_srgb(integral / 106.9461715), 1.0, atol=1.5e-2)


def textbook_xyz_to_srgb():
    bradford = np.array([[0.8951, 0.2664, -0.1614], [-0.7502, 1.7135, 0.0367], [0.0389, -0.0685, 1.0296]])
    e_white = np.array([1.0, 1.0, 1.0])
    d65 = np.array([0.95047, 1.0, 1.08883])
    adapt = np.linalg.inv(bradford) @ np.diag((bradford @ d65) / (bradford @ e_white)) @ bradford
    srgb = np.array([[3.2406, -1.5372, -0.4986], [-0.9689, 1.8758, 0.0415], [0.0557, -0.2040, 1.0570]])
    return srgb @ adapt


def test_xyz_to_srgb_is_bradford_then_iec_srgb():
    m = textbook_xyz_to_srgb()
    rng = np.random.default_rng(8)
    for xyz in rng.uniform(0.0, 2.0, (50, 3)):
        assert np.allclose(O.xyz_to_srgb(xyz), m @ xyz, atol=2e-3 * max(1.0, xyz.max()))


@pytest.fixture(scope="module")
def lut_buffers():
    r = cpu_renderer()
    b = r.export()
    assert b.rgb2specFloatCount > 0
    yield b
    r.close()


def test_rgb2spec_interval_search_forms_agree(lut_buffers):
    """The shading kernel finds a colour's brightness interval with two levels of
    eight-way comparisons; the oracle (and the reference) with a binary search.
    Both must give the same index on the shipped table for every brightness,
    including the table's own entries and their float neighbours."""
    table = np.ctypeslib.as_array(lut_buffers.rgb2specTable, (lut_buffers.rgb2specFloatCount,))
    res = 64  # the payload is res scale values, then 3 * res^3 * 3 coefficients
    assert len(table) == res + 9 * res ** 3
    sc = table[:res].astype(np.float32)
    assert np.all(np.diff(sc) >= 0)

    def binary(x):
        left, size = 0, res - 2
        while size > 0:
            half = size >> 1
            middle = left + half + 1
            if sc[middle] <= x:
                left, size = middle, size - half - 1
            else:
                size = half
        return min(left, res - 2)

    def eight_way(x):
        b = sum(int(sc[8 * k] <= x) for k in range(1, 8))
        return min(8 * b + sum(int(sc[8 * b + k] <= x) for k in range(1, 8)), res - 2)

    xs = np.concatenate([np.random.default_rng(3).random(20000).astype(np.float32), sc,
                         np.nextafter(sc, np.float32(2)), np.nextafter(sc, np.float32(-1)),
                         np.float32([0.0, 1.0, 1e-8, 4.0])])
    assert all(binary(x) == eight_way(x) for x in xs)


def test_rgb2spec_round_trip(lut_buffers):
    """sRGB -> coefficients -> spectrum -> CIE XYZ -> sRGB returns the input."""
    lam = np.arange(360.0, 831.0)
    cmf = np.array([O.cie_xyz(float(l)) for l in lam])
    w = np.ones_like(lam)
    w[0] = w[-1] = 0.5
    rng = np.random.default_rng(12)
    worst = 0.0
    for rgb in rng.uniform(0.05, 0.95, (24, 3)):
        s = np.array([O.rgb2spec(lut_buffers, rgb, float(l)) for l in lam])
        assert np.all((s > 0) & (s < 1))
        xyz = (s[:, None] * cmf * w[:, None]).sum(0) / 106.9461715
        back = O.xyz_to_srgb(xyz)
        worst = max(worst, float(np.abs(back - rgb).max()))
    assert worst < 5e-3


def test_rgb2spec_grey_is_nearly_flat(lut_buffers):
    # the fitted CMFs integrate to slightly different totals, so the exact
    # solution for grey is flat only to about 2%
    for g in (0.1, 0.5, 0.9):
        s = [O.rgb2spec(lut_buffers, [g, g, g], float(l)) for l in (400.0, 500.0, 600.0, 700.0)]
        assert np.allclose(s, g, rtol=3e-2)


def test_hero_wavelengths_are_stratified():
    for hs in (1, 77, 123456789):
        for si in range(8):
            lam = O.hero_wavelengths(hs, si)
            assert np.all((lam >= 360.0) & (lam < 830.0))
            steps = np.mod(np.diff(np.concatenate([lam, lam[:1]])), 470.0)
            assert np.allclose(steps, 117.5, atol=1e-6)


def test_dispersive_ior_follows_the_abbe_number():
    assert O.dispersive_ior(1.5, 0.0, 450.0) == pytest.approx(1.5)
    assert O.dispersive_ior(1.5, 40.0, 587.56) == pytest.approx(1.5, abs=1e-6)
    nf, nc = O.dispersive_ior(1.5, 40.0, 486.13), O.dispersive_ior(1.5, 40.0, 656.27)
    assert (1.5 - 1.0) / (nf - nc) == pytest.approx(40.0, rel=1e-3)


# -------------------------------------------------------------- Fresnel --
def textbook_fresnel(ci, eta):
    s2 = max(0.0, 1.0 - ci * ci) / (eta * eta)
    if s2 >= 1.0:
        return 1.0
    ct = math.sqrt(1.0 - s2)
    rs = (ci - eta * ct) / (ci + eta * ct)
    rp = (eta * ci - ct) / (eta * ci + ct)
    return 0.5 * (rs * rs + rp * rp)


def test_fresnel_dielectric_textbook():
    for eta in (1.5, 1.33, 2.4):
        assert O.fresnel_dielectric(1.0, eta) == pytest.approx(((eta - 1) / (eta + 1)) ** 2, rel=1e-9)
        assert O.fresnel_dielectric(1.0, 1.0 / eta) == pytest.approx(((eta - 1) / (eta + 1)) ** 2, rel=1e-9)
        for ci in np.linspace(0.01, 1.0, 40):
            assert O.fresnel_dielectric(ci, eta) == pytest.approx(textbook_fresnel(ci, eta), abs=1e-9)
            assert O.fresnel_dielectric(ci, 1 / eta) == pytest.approx(textbook_fresnel(ci, 1 / eta), abs=1e-9)
    crit = math.sqrt(1.0 - (1 / 1.5) ** 2)
    assert O.fresnel_dielectric(crit * 0.98, 1 / 1.5) == 1.0  # total internal reflection


# ----------------------------------------------------------------- BSDF --
def local_dir(theta_deg, phi_deg=0.0):
    t, p = math.radians(theta_deg), math.radians(phi_deg)
    return np.array([math.sin(t) * math.cos(p), math.sin(t) * math.sin(p), math.cos(t)])


LAMBERT = dict(roughness=1.0, specular=0.0, metallic=0.0)


def test_lambert_closed_form(lut_buffers):
    rho = (0.7, 0.4, 0.2)
    m = V.default_material(baseColor=rho, **LAMBERT)
    wo = local_dir(30, 10)
    for th, ph in ((0, 0), (25, 80), (60, 200), (85, 300)):
        wi = local_dir(th, ph)
        val, pdf = O.bsdf_eval(lut_buffers, m, wo, wi)
        # eval returns f (without the cosine) and the cosine-hemisphere pdf
        assert pdf == pytest.approx(math.cos(math.radians(th)) / math.pi, rel=1e-9)
        assert np.allclose(val, np.float32(rho) / math.pi, rtol=1e-9)
        # single-wavelength mode lifts the albedo through rgb2spec
        v1, p1 = O.bsdf_eval(lut_buffers, m, wo, wi, wavelength=550.0, mode=1)
        assert v1[0] == pytest.approx(O.rgb2spec(lut_buffers, rho, 550.0) / math.pi, rel=1e-6)
    val, pdf = O.bsdf_eval(lut_buffers, m, wo, local_dir(120))
    assert pdf == 0.0 and np.all(val == 0.0)
    s = O.bsdf_sample_batch(lut_buffers, m, wo, 2000, seed=3)
    assert np.allclose(s[:, 3:6], np.float32(rho), rtol=1e-9)  # weight = the (fp32) albedo exactly


MATERIALS = {
    "coated_diffuse": dict(baseColor=(0.6, 0.5, 0.4), roughness=0.4, metallic=0.0, specular=0.5),
    "rough_metal": dict(baseColor=(0.9, 0.6, 0.3), roughness=0.35, metallic=1.0),
    "smooth_metal": dict(baseColor=(0.9, 0.9, 0.9), roughness=0.1, metallic=1.0),
    "clearcoat": dict(baseColor=(0.3, 0.5, 0.7), roughness=0.6, clearcoat=1.0, clearcoatGloss=0.7),
    "rough_glass": dict(baseColor=(1.0, 1.0, 1.0), roughness=0.4, transmission=1.0, ior=1.5),
    "sheen_diffuse": dict(baseColor=(0.5, 0.2, 0.2), roughness=1.0, specular=0.0, sheenTintWeight=(1.0, 0.9, 0.8, 1.0),
                          sheenRoughness=0.5),
    "sheen_clearcoat": dict(baseColor=(0.2, 0.3, 0.6), roughness=0.5, sheenTintWeight=(0.6, 0.6, 1.0, 0.7),
                            sheenRoughness=0.2, clearcoat=0.5, clearcoatGloss=0.6),
}


def sphere_dirs(n):
    i = np.arange(n) + 0.5
    z = 1.0 - 2.0 * i / n
    phi = i * math.pi * (3.0 - math.sqrt(5.0))
    r = np.sqrt(1.0 - z * z)
    return np.stack([r * np.cos(phi), r * np.sin(phi), z], 1)


@pytest.mark.parametrize("name", sorted(MATERIALS))
def test_bsdf_sampling_agrees_with_eval(lut_buffers, name):
    m = V.default_material(**MATERIALS[name])
    wo = local_dir(40, 25)
    s = O.bsdf_sample_batch(lut_buffers, m, wo, 3000, seed=11)
    usable = s[:, 6] > 0
    assert usable.mean() > 0.5
    for row in s[usable][:400]:
        val, pdf = O.bsdf_eval(lut_buffers, m, wo, row[:3])
        assert pdf == pytest.approx(row[6], rel=2e-4, abs=1e-9)
        assert np.allclose(val * abs(row[2]) / pdf, row[3:6], rtol=2e-4, atol=1e-7)
    # the eval pdf integrates (over the sphere) to the probability of a usable sample
    dirs = sphere_dirs(40000)
    pdfs = np.array([O.bsdf_eval(lut_buffers, m, wo, d)[1] for d in dirs])
    integral = pdfs.mean() * 4.0 * math.pi
    assert integral <= 1.0 + 0.03
    assert integral == pytest.approx(usable.mean(), abs=0.04)


# ---------------------------------------------------------------- sheen --
SHEEN_REFERENCE_TABLE = "/root/reference/src/shaders/bsdf/data/sheen_ltc.slang"


def carried_sheen_table():
    import re
    text = open(os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                             "vkrt_amd", "csrc", "sheen_ltc_table.inc")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return np.array([float(v) for v in re.findall(r"-?\d+\.\d+", text)], np.float32).reshape(3, 32, 32)


@pytest.mark.skipif(not os.path.exists(SHEEN_REFERENCE_TABLE), reason="reference sources are not mounted")
def test_sheen_table_is_the_reference_table():
    import re
    text = open(SHEEN_REFERENCE_TABLE).read()
    body = text[text.index("SHEEN_LTC_TABLE[3072]"):]
    body = body[body.index("{") + 1:body.index("}")]
    ref = np.array([float(v) for v in re.findall(r"-?\d+\.\d+(?:e-?\d+)?", body)], np.float32)
    np.testing.assert_array_equal(carried_sheen_table().ravel(), ref)


def test_sheen_lookup_hits_table_nodes_and_interpolates():
    """sheen.slang:22-39: rows are roughness, columns cos(theta), layers a, b,
    albedo; exact at the nodes, bilinear between them, clamped outside."""
    t = carried_sheen_table()
    assert t[2].min() >= 0.0 and t[2].max() <= 1.0  # a directional albedo
    for layer in range(3):
        for i, j in ((0, 0), (5, 17), (31, 31), (12, 3), (20, 30)):
            assert O.sheen_ltc(j / 31.0, i / 31.0, layer) == pytest.approx(float(t[layer, i, j]), abs=1e-7)
        c, r = (7.25) / 31.0, (9.5) / 31.0
        want = ((t[layer, 9, 7] * 0.75 + t[layer, 9, 8] * 0.25) * 0.5 + (t[layer, 10, 7] * 0.75 + t[layer, 10, 8] * 0.25) * 0.5)
        assert O.sheen_ltc(c, r, layer) == pytest.approx(float(want), rel=1e-5, abs=1e-7)
        assert O.sheen_ltc(1.7, -0.3, layer) == pytest.approx(float(t[layer, 0, 31]), abs=1e-7)


@pytest.mark.parametrize("theta,rough", [(10, 0.3), (45, 0.5), (70, 0.8), (30, 1.0)])
def test_sheen_density_is_a_normalised_ltc(theta, rough):
    """The lobe's density is a linearly transformed clamped cosine, which
    integrates to 1 over the upper hemisphere whatever the transform; the
    value is the directional albedo times it, so the lobe's energy is the
    table's albedo (at most 1)."""
    wo = local_dir(theta, 30)
    dirs = sphere_dirs(200000)
    dirs = dirs[dirs[:, 2] > 0]
    vals = np.array([O.sheen_eval(rough, wo, d) for d in dirs])
    integral = vals[:, 1].mean() * 2.0 * math.pi
    assert integral == pytest.approx(1.0, abs=0.01)
    albedo = O.sheen_ltc(math.cos(math.radians(theta)), rough, 2)
    np.testing.assert_allclose(vals[:, 0], albedo * vals[:, 1], rtol=1e-12, atol=0)
    assert 0.0 < albedo <= 1.0
    # below the horizon on either side there is no sheen
    assert O.sheen_eval(rough, wo, local_dir(120)) == (0.0, 0.0)
    assert O.sheen_eval(rough, local_dir(100), local_dir(20)) == (0.0, 0.0)


def test_sheen_attenuates_the_base_by_its_albedo(lut_buffers):
    """weights.slang:27-62 / eval_rgb.slang:92-100: a sheen layer adds its lobe
    and scales everything under it by 1 - weight * albedo(wo)."""
    base = dict(baseColor=(0.6, 0.6, 0.6), **LAMBERT)
    wo, wi = local_dir(35, 0), local_dir(50, 140)
    v0, _ = O.bsdf_eval(lut_buffers, V.default_material(**base), wo, wi)
    rough, w = 0.4, 0.8
    m = V.default_material(sheenTintWeight=(1.0, 1.0, 1.0, w), sheenRoughness=rough, **base)
    v1, _ = O.bsdf_eval(lut_buffers, m, wo, wi)
    att = 1.0 - w * O.sheen_ltc(wo[2], rough, 2)
    lobe = w * O.sheen_eval(rough, wo, wi)[0]
    np.testing.assert_allclose(v1, v0 * att + lobe, rtol=1e-6)


def test_metal_reciprocity_and_energy(lut_buffers):
    m = V.default_material(baseColor=(1.0, 1.0, 1.0), roughness=0.3, metallic=1.0)
    for a, b in (((20, 0), (50, 170)), ((70, 30), (10, 250)), ((45, 90), (45, 270))):
        wo, wi = local_dir(*a), local_dir(*b)
        f1 = O.bsdf_eval(lut_buffers, m, wo, wi)[0]
        f2 = O.bsdf_eval(lut_buffers, m, wi, wo)[0]
        assert np.allclose(f1, f2, rtol=1e-9)
    s = O.bsdf_sample_batch(lut_buffers, m, local_dir(30), 20000, seed=2)
    albedo = s[:, 3].sum() / len(s)
    assert 0.85 < albedo <= 1.0 + 5e-3


# ---------------------------------------------------------- environment --
def test_environment_latlong_mapping():
    r = cpu_renderer()
    w, h = 64, 32
    tex = np.zeros((h, w, 4), np.float32)
    tex[..., 0] = np.arange(w)[None, :]
    tex[..., 1] = np.arange(h)[:, None]
    tex[..., 3] = 1.0
    r.set_environment_texture(r.add_texture_rgba32f(tex))
    r.set_environment_light((1.0, 1.0, 1.0), 2.0)
    b = r.export()
    for x, y in ((0, 0), (5, 9), (31, 16), (63, 31), (40, 3)):
        u, v = (x + 0.5) / w, (y + 0.5) / h
        phi, theta = (u - 0.5) * 2 * math.pi, v * math.pi
        d = [math.sin(theta) * math.cos(phi), math.sin(theta) * math.sin(phi), math.cos(theta)]
        assert np.allclose(O.environment(b, d), (2.0 * x, 2.0 * y, 0.0), atol=2e-3)
    r.close()


# ----------------------------------------------------------- integrator --
def furnace_scene(mode, spectral=V.SPECTRAL_HERO):
    r = cpu_renderer(16, 16)
    v, i = faceted_sphere(6)
    mat = r.add_material(V.default_material(baseColor=(1.0, 1.0, 1.0), **LAMBERT))
    r.set_mesh_material(r.add_mesh(v, i), mat)
    r.set_camera((0.0, -3.0, 0.0), (0.0, 0.0, 0.0), (0.0, 0.0, 1.0), 30.0)
    r.set_environment_light((1.0, 1.0, 1.0), 1.0)
    r.configure(spp=8, max_depth=12, min_depth=4, mode=mode, spectral=spectral, rng=V.RNG_PCG32, seed=1)
    return r


def test_white_furnace_rgb_is_exact():
    r = furnace_scene(V.RENDER_MODE_RGB)
    osc = O.OracleScene(r.export())
    acc, _ = osc.render(16, 16, precision=64)
    assert np.allclose(acc[..., :3], 1.0, atol=1e-9)
    assert np.all(acc[..., 3] == 8)
    osc.close()
    r.close()


@pytest.mark.parametrize("spectral", [V.SPECTRAL_HERO, V.SPECTRAL_SINGLE])
def test_white_furnace_spectral(spectral):
    r = furnace_scene(V.RENDER_MODE_SPECTRAL, spectral)
    osc = O.OracleScene(r.export())
    acc, _ = osc.render(16, 16, precision=64)
    # accumulation holds CIE XYZ of an equal-energy-normalised spectrum
    rgb = np.einsum("ij,hwj->hwi", textbook_xyz_to_srgb(), acc[..., :3])
    assert abs(rgb.mean() - 1.0) < 2e-2
    osc.close()
    r.close()


def test_closed_emissive_sphere_radiance():
    """Inside a closed sphere whose walls emit Le and reflect rho (Lambert),
    radiance is Le / (1 - rho) everywhere; NEE, MIS and Russian roulette must
    leave the expectation unchanged."""
    r = cpu_renderer(24, 24)
    v, i = faceted_sphere(4, radius=2.0, inward=True)
    mat = r.add_material(V.default_material(baseColor=(0.5, 0.5, 0.5), emissionColor=(1.0, 1.0, 1.0),
                                            emissionLuminance=1.0, **LAMBERT))
    r.set_mesh_material(r.add_mesh(v, i), mat)
    r.set_camera((0.1, -0.3, 0.2), (0.0, 1.0, 0.0), (0.0, 0.0, 1.0), 60.0)
    r.set_environment_light((0.0, 0.0, 0.0), 0.0)
    r.configure(spp=64, max_depth=64, min_depth=4, mode=V.RENDER_MODE_RGB, rng=V.RNG_PCG32, seed=1)
    osc = O.OracleScene(r.export())
    acc, _ = osc.render(24, 24, precision=64)
    px = acc[..., :3].reshape(-1)
    se = px.std() / math.sqrt(len(px) / 3)
    assert abs(px.mean() - 2.0) < 4 * se + 1e-3, (px.mean(), se)
    osc.close()
    r.close()


def test_accumulation_is_the_running_mean():
    r = furnace_scene(V.RENDER_MODE_RGB)
    r.set_camera((0.0, -3.0, 0.0), (0.4, 0.0, 0.0), (0.0, 0.0, 1.0), 50.0)
    r.configure(spp=3)
    osc = O.OracleScene(r.export())
    one = osc.render(16, 16, precision=64)[0]
    two = osc.render(16, 16, accum=one.copy(), precision=64)[0]
    assert np.all(two[..., 3] == 6)
    # same frame number twice -> identical samples -> unchanged mean
    assert np.allclose(two[..., :3], one[..., :3], atol=1e-12)
    osc.close()
    r.close()


# ------------------------------------------------- environment sampling --
SKY = dict(width=256, height=128, sun_elevation_deg=50.0, sun_radiance=50.0, sun_diameter_deg=6.0)


@pytest.fixture(scope="module")
def sky_buffers():
    r = cpu_renderer(8, 8)
    r.set_environment_texture(r.add_texture_rgba32f(scenes.sky_environment(**SKY)))
    r.set_environment_light((1.0, 1.0, 1.0), 1.0)
    r.configure(env_importance=True)
    yield r, r.export()
    r.close()


def test_environment_table_matches_luminance_times_sine(sky_buffers):
    _, b = sky_buffers
    w, h = b.environmentWidth, b.environmentHeight
    n = w * h
    pix = np.ctypeslib.as_array(b.environmentPixels, shape=(h, w, 4)).astype(np.float64)
    lum = np.maximum(pix[..., :3] @ np.array([0.2126, 0.7152, 0.0722]), 0.0)
    sin_t = np.sin(np.pi * (np.arange(h) + 0.5) / h)[:, None]
    weight = lum * sin_t
    weight = weight + 1e-3 * weight.sum() / n * sin_t  # floor keeps every direction reachable
    pmf = (weight / weight.sum()).reshape(-1)
    got = np.ctypeslib.as_array(b.environmentPdf, shape=(n,))
    assert np.allclose(got, pmf, rtol=1e-5, atol=1e-12)
    q = np.ctypeslib.as_array(b.environmentAliasQ, shape=(n,))
    idx = np.ctypeslib.as_array(b.environmentAliasIdx, shape=(n,))
    assert np.allclose(alias_pmf(q, idx), pmf, rtol=1e-4, atol=1e-7)  # fp32 table entries


def test_environment_sampler_density(sky_buffers):
    """Sampled directions carry their solid-angle density: the pdf recomputed
    from the direction agrees, and Monte Carlo integrals of known functions are
    recovered (4 pi for 1, pi for the clamped cosine)."""
    _, b = sky_buffers
    n = 200000
    s = O.env_sample_batch(b, n, seed=5)
    assert np.all(s[:, 5] == 1.0)
    assert np.allclose(s[:, 3], s[:, 4], rtol=1e-6)
    assert np.allclose(np.linalg.norm(s[:, :3], axis=1), 1.0, atol=1e-12)
    for g, exact in ((np.ones(n), 4 * math.pi), (np.maximum(s[:, 2], 0.0), math.pi)):
        v = g / s[:, 3]
        assert abs(v.mean() - exact) < 4 * v.std() / math.sqrt(n)


def flat_quad_under_sky(rho=0.5):
    r = cpu_renderer(12, 12)
    pos = np.array([[-5, -5, 0], [5, -5, 0], [5, 5, 0], [-5, 5, 0]], np.float32)
    verts = V.make_vertices(pos, np.tile([[0, 0, 1]], (4, 1)).astype(np.float32))
    mat = r.add_material(V.default_material(baseColor=(rho,) * 3, **LAMBERT))
    r.set_mesh_material(r.add_mesh(verts, np.array([[0, 1, 2], [0, 2, 3]], np.uint32)), mat)
    r.set_camera((0.0, -0.01, 3.0), (0.0, 0.0, 0.0), (0.0, 1.0, 0.0), 20.0)
    r.set_environment_texture(r.add_texture_rgba32f(scenes.sky_environment(**SKY)))
    r.set_environment_light((1.0, 1.0, 1.0), 1.0)
    return r


def test_environment_nee_matches_the_irradiance_integral():
    """A Lambertian plane under the sky reflects rho/pi * integral of L cos over
    the upper hemisphere; both the reference estimator (environment gathered on
    BSDF misses) and environment NEE with MIS must reproduce it."""
    rho = 0.5
    r = flat_quad_under_sky(rho)
    b = r.export()
    W, H = 512, 256  # midpoint quadrature over the upper hemisphere
    ph = ((np.arange(W) + 0.5) / W - 0.5) * 2 * math.pi
    total = np.zeros(3)
    for j in range(H // 2):
        th = (j + 0.5) / H * math.pi
        st, ct = math.sin(th), math.cos(th)
        dirs = np.stack([st * np.cos(ph), st * np.sin(ph), np.full(W, ct)], 1)
        total += np.array([O.environment(b, d) for d in dirs]).sum(0) * ct * st * (2 * math.pi / W) * (math.pi / H)
    exact = rho / math.pi * total
    for nee, spp in ((False, 2048), (True, 512)):
        r.configure(spp=spp, max_depth=2, min_depth=4, mode=V.RENDER_MODE_RGB, rng=V.RNG_PCG32, seed=7,
                    env_importance=nee)
        osc = O.OracleScene(r.export())
        acc, _ = osc.render(12, 12, precision=64)
        osc.close()
        px = acc[..., :3].reshape(-1, 3)
        se = px.std(0) / math.sqrt(len(px))
        assert np.all(np.abs(px.mean(0) - exact) < 4 * se + 2e-3 * exact), (nee, px.mean(0), exact, se)
    r.close()
