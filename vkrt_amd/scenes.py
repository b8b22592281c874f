"""Procedural scenes for the benchmark configurations.

cornell_box   -- 5 wall quads + two closed axis-aligned boxes (heights 0.3 and
                 0.6) = 34 triangles, plus a 2-triangle ceiling light of
                 radiance 15; Lambertian materials expressed in the reference's
                 principled Material (roughness 1, specular 0, metallic 0 makes
                 its BSDF exactly Lambertian).
displaced_sphere -- geodesic icosphere of frequency f (20*f^2 triangles; f = 229
                 gives 1,048,820), displaced along the normal by seeded
                 3-octave value noise of amplitude 0.1*radius, split into 64
                 meshes by material id = hash(face) mod 64.
sky_environment -- lat-long RGBA32F sky gradient with a 0.5 degree sun disc.
stress_scene  -- 16 instances of the displaced sphere under seeded rigid
                 transforms plus 4M random small triangles (edge U(1e-3, 1e-2))
                 in the unit cube, all diffuse: about 20.8M triangles.
Scenes are returned as plain data and loaded through the C API by `load`.
"""

from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from . import Renderer, default_material, make_vertices


@dataclass
class MeshData:
    vertices: np.ndarray  # N x 20 float32 (reference Vertex layout)
    indices: np.ndarray  # M x 3 uint32
    material: int  # index into Scene.materials
    transform: np.ndarray | None = None  # 4 x 4 world transform (row-major), identity if None
    opacity: float = 1.0  # mesh opacity (VKRT_setMeshOpacity)


@dataclass
class Scene:
    meshes: list = field(default_factory=list)
    materials: list = field(default_factory=list)
    textures: list = field(default_factory=list)  # (H x W x 4 pixels, format, colour space), added first
    camera: tuple = ((0.0, -1.0, 0.0), (0.0, 0.0, 0.0), (0.0, 0.0, 1.0), 40.0)
    environment: np.ndarray | None = None  # H x W x 4 float32
    environment_color: tuple = (0.0, 0.0, 0.0)
    environment_strength: float = 0.0

    @property
    def triangle_count(self) -> int:
        return sum(len(m.indices) for m in self.meshes)


def _quad(p0, p1, p2, p3, normal):
    pos = np.array([p0, p1, p2, p3], dtype=np.float32)
    nrm = np.tile(np.asarray(normal, dtype=np.float32), (4, 1))
    return pos, nrm, np.array([[0, 1, 2], [0, 2, 3]], dtype=np.uint32)


def _merge(parts):
    pos, nrm, idx = [], [], []
    base = 0
    for p, n, i in parts:
        pos.append(p)
        nrm.append(n)
        idx.append(i + base)
        base += len(p)
    return make_vertices(np.concatenate(pos), np.concatenate(nrm)), np.concatenate(idx).astype(np.uint32)


def _box(x0, x1, y0, y1, h):
    """Closed axis-aligned box on the floor: 6 faces, outward normals."""
    z0, z1 = 0.0, h
    return [
        _quad((x0, y0, z1), (x1, y0, z1), (x1, y1, z1), (x0, y1, z1), (0, 0, 1)),
        _quad((x0, y0, z0), (x0, y1, z0), (x1, y1, z0), (x1, y0, z0), (0, 0, -1)),
        _quad((x0, y0, z0), (x1, y0, z0), (x1, y0, z1), (x0, y0, z1), (0, -1, 0)),
        _quad((x1, y1, z0), (x0, y1, z0), (x0, y1, z1), (x1, y1, z1), (0, 1, 0)),
        _quad((x0, y1, z0), (x0, y0, z0), (x0, y0, z1), (x0, y1, z1), (-1, 0, 0)),
        _quad((x1, y0, z0), (x1, y1, z0), (x1, y1, z1), (x1, y0, z1), (1, 0, 0)),
    ]


def lambert(rgb, emission=0.0):
    m = default_material(baseColor=rgb, roughness=1.0, specular=0.0, metallic=0.0)
    m.emissionLuminance = float(emission)
    return m


def cornell_box() -> Scene:
    """Unit-cube Cornell box, z up, open towards -y, camera on the -y side."""
    white, red, green = (0.73, 0.73, 0.73), (0.63, 0.065, 0.05), (0.14, 0.45, 0.09)
    s = Scene()
    s.materials = [lambert(white), lambert(red), lambert(green), lambert(white, emission=15.0)]
    walls = [
        _quad((0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0), (0, 0, 1)),  # floor
        _quad((0, 0, 1), (0, 1, 1), (1, 1, 1), (1, 0, 1), (0, 0, -1)),  # ceiling
        _quad((0, 1, 0), (1, 1, 0), (1, 1, 1), (0, 1, 1), (0, -1, 0)),  # back
    ]
    s.meshes.append(MeshData(*_merge(walls), 0))
    s.meshes.append(MeshData(*_merge([_quad((0, 0, 0), (0, 1, 0), (0, 1, 1), (0, 0, 1), (1, 0, 0))]), 1))
    s.meshes.append(MeshData(*_merge([_quad((1, 0, 0), (1, 0, 1), (1, 1, 1), (1, 1, 0), (-1, 0, 0))]), 2))
    s.meshes.append(MeshData(*_merge(_box(0.53, 0.83, 0.12, 0.42, 0.3)), 0))
    s.meshes.append(MeshData(*_merge(_box(0.17, 0.47, 0.48, 0.78, 0.6)), 0))
    zl = 1.0 - 1e-3
    light = _quad((0.37, 0.37, zl), (0.63, 0.37, zl), (0.63, 0.63, zl), (0.37, 0.63, zl), (0, 0, -1))
    s.meshes.append(MeshData(*_merge([light]), 3))
    s.camera = ((0.5, -1.35, 0.5), (0.5, 0.5, 0.5), (0.0, 0.0, 1.0), 40.0)
    s.environment_color = (0.0, 0.0, 0.0)
    s.environment_strength = 0.0
    return s


def sheen_cornell() -> Scene:
    """The Cornell box with microfiber sheen on both boxes
    (src/shaders/bsdf/lobes/sheen.slang): a velvet-like tall box (white tint,
    full weight, sheen roughness 0.3 over a dark red base) and a short box with
    a tinted half-weight sheen (roughness 0.8) over a clearcoated blue base."""
    s = cornell_box()
    velvet = default_material(baseColor=(0.35, 0.05, 0.05), roughness=0.9, metallic=0.0, specular=0.3)
    velvet.sheenTintWeight[:] = (1.0, 1.0, 1.0, 1.0)
    velvet.sheenRoughness = 0.3
    satin = default_material(baseColor=(0.1, 0.2, 0.6), roughness=0.4, metallic=0.0, specular=0.5, clearcoat=0.4,
                             clearcoatGloss=0.7)
    satin.sheenTintWeight[:] = (0.9, 0.8, 0.5, 0.5)
    satin.sheenRoughness = 0.8
    s.materials += [velvet, satin]
    s.meshes[4].material = len(s.materials) - 2  # tall box
    s.meshes[3].material = len(s.materials) - 1  # short box
    return s


def alpha_cornell() -> Scene:
    """The Cornell box with three surfaces that run the alpha test
    (src/shaders/rt/alpha_test.slang): a horizontal BLEND sheet of opacity 0.5
    under the light (shadow rays cross it), a MASK quad before the back wall
    whose vertex-colour alpha runs 0 -> 1 from left to right (cutoff 0.4), and
    the red wall drawn at mesh opacity 0.35 (not a box: a box's bottom face is
    coplanar with the floor, and rays through a translucent box would meet two
    surfaces at the same distance)."""
    s = cornell_box()
    sheet = default_material(baseColor=(0.9, 0.85, 0.6), roughness=0.6, metallic=0.0, specular=0.0)
    sheet.alphaMode = 2  # BLEND
    sheet.opacity = 0.5
    mask = default_material(baseColor=(0.2, 0.3, 0.8), roughness=1.0, metallic=0.0, specular=0.0)
    mask.alphaMode = 1  # MASK
    mask.alphaCutoff = 0.4
    s.materials += [sheet, mask]
    zs = 0.8
    pos, nrm, idx = _quad((0.25, 0.2, zs), (0.75, 0.2, zs), (0.75, 0.8, zs), (0.25, 0.8, zs), (0, 0, -1))
    s.meshes.append(MeshData(make_vertices(pos, nrm), idx, len(s.materials) - 2))
    pos, nrm, idx = _quad((0.1, 0.9, 0.1), (0.9, 0.9, 0.1), (0.9, 0.9, 0.7), (0.1, 0.9, 0.7), (0, -1, 0))
    colors = np.ones((4, 4), dtype=np.float32)
    colors[:, 3] = [0.0, 1.0, 1.0, 0.0]
    s.meshes.append(MeshData(make_vertices(pos, nrm, colors), idx, len(s.materials) - 1))
    s.meshes[1].opacity = 0.35  # the red wall
    return s


def _wrap(u, v):
    return int(u) | (int(v) << 16)


def _textured_quad(corners, normal, uv_scale1=2.0):
    pos, nrm, idx = _quad(*corners, normal)
    uv = np.array([[0, 0], [1, 0], [1, 1], [0, 1]], dtype=np.float32)
    return make_vertices(pos, nrm, texcoord0=uv, texcoord1=uv * uv_scale1), idx


def textured_cornell() -> Scene:
    """A Cornell box whose materials sample every kind of material texture
    (src/shaders/material/textures.slang): an sRGB checker on the floor with a
    scaled, rotated and offset UV transform; a normal map and a 16-bit
    metallic-roughness map on a glossy back wall (mirrored / clamped wrap); a
    half-float base colour on the red wall read through texture-coordinate
    set 1; an emissive panel with a mirrored-repeat emissive map; a MASK quad
    cut by its base-colour texture's alpha and a BLEND sheet whose opacity is
    that alpha. Texture-coordinate set 0 runs 0 -> 1 across every quad, set 1
    twice as fast."""
    import vkrt_amd as V
    n = 32
    y, x = np.mgrid[0:n, 0:n]
    checker = np.zeros((n, n, 4), np.uint8)
    on = ((x // 4 + y // 4) % 2) == 1
    checker[..., 0] = np.where(on, 230, 40)
    checker[..., 1] = np.where(on, 200, 90 + 4 * x)
    checker[..., 2] = np.where(on, 60, 50 + 5 * y)
    checker[..., 3] = 255
    bump = np.zeros((n, n, 4), np.uint8)  # tangent-space normals of a sinusoidal height field
    gx = 0.6 * np.cos(2 * np.pi * x / n * 3)
    gy = 0.6 * np.sin(2 * np.pi * y / n * 2)
    nv = np.stack([-gx, -gy, np.ones_like(gx)], -1)
    nv /= np.linalg.norm(nv, axis=-1, keepdims=True)
    bump[..., :3] = np.round((nv * 0.5 + 0.5) * 255).astype(np.uint8)
    bump[..., 3] = 255
    m = 16
    yy, xx = np.mgrid[0:m, 0:m]
    mr = np.zeros((m, m, 4), np.uint16)
    mr[..., 1] = (65535 * (0.15 + 0.85 * xx / (m - 1))).astype(np.uint16)  # roughness factor
    mr[..., 2] = np.where((yy // 4) % 2 == 0, 65535, 6000).astype(np.uint16)  # metallic factor
    mr[..., 3] = 65535
    red = np.zeros((8, 8, 4), np.float16)
    gy8, gx8 = np.mgrid[0:8, 0:8]
    red[..., 0] = 0.5 + 0.5 * (gx8 / 7)
    red[..., 1] = 0.2 + 0.6 * ((gx8 + gy8) % 2)
    red[..., 2] = 0.3 * gy8 / 7
    red[..., 3] = 1.0
    emis = np.zeros((8, 8, 4), np.uint8)
    emis[..., 0] = 255
    emis[..., 1] = (gx8 * 36).astype(np.uint8)
    emis[..., 2] = ((gx8 + gy8) % 2 * 255).astype(np.uint8)
    emis[..., 3] = 255
    holes = np.full((m, m, 4), 255, np.uint8)
    cx = (xx % 8) - 3.5
    cy = (yy % 8) - 3.5
    holes[..., 3] = np.where(cx * cx + cy * cy < 7.0, 0, np.clip(40 + 13 * xx, 0, 255)).astype(np.uint8)
    holes[..., 0] = 90
    holes[..., 1] = 140
    holes[..., 2] = 220
    srgb, lin = V.TEXTURE_COLOR_SPACE_SRGB, V.TEXTURE_COLOR_SPACE_LINEAR
    s = Scene()
    s.textures = [(checker, V.TEXTURE_FORMAT_RGBA8_UNORM, srgb), (bump, V.TEXTURE_FORMAT_RGBA8_UNORM, lin),
                  (mr, V.TEXTURE_FORMAT_RGBA16_UNORM, lin), (red, V.TEXTURE_FORMAT_RGBA16_SFLOAT, lin),
                  (emis, V.TEXTURE_FORMAT_RGBA8_UNORM, srgb), (holes, V.TEXTURE_FORMAT_RGBA8_UNORM, srgb)]
    floor = lambert((0.9, 0.9, 0.9))
    floor.baseColorTextureIndex = 0
    floor.baseColorTextureTransform[:] = (3.0, 3.0, 0.1, 0.2)
    floor.textureRotations[0] = 0.4
    back = default_material(baseColor=(0.8, 0.8, 0.75), roughness=0.6, metallic=0.9, specular=0.5)
    back.normalTextureIndex = 1
    back.normalTextureScale = 0.8
    back.normalTextureWrap = _wrap(V.TEXTURE_WRAP_MIRRORED_REPEAT, V.TEXTURE_WRAP_CLAMP_TO_EDGE)
    back.normalTextureTransform[:] = (2.5, 1.5, -0.3, 0.0)
    back.metallicRoughnessTextureIndex = 2
    back.metallicRoughnessTextureWrap = _wrap(V.TEXTURE_WRAP_CLAMP_TO_EDGE, V.TEXTURE_WRAP_CLAMP_TO_EDGE)
    back.metallicRoughnessTextureTransform[:] = (1.4, 1.4, -0.2, -0.2)
    redm = lambert((1.0, 0.3, 0.3))
    redm.baseColorTextureIndex = 3
    redm.textureTexcoordSets = 1 << (8 * 0)  # base colour from set 1
    redm.baseColorTextureWrap = _wrap(V.TEXTURE_WRAP_MIRRORED_REPEAT, V.TEXTURE_WRAP_REPEAT)
    panel = lambert((0.2, 0.2, 0.2), emission=4.0)
    panel.emissiveTextureIndex = 4
    panel.emissiveTextureWrap = _wrap(V.TEXTURE_WRAP_MIRRORED_REPEAT, V.TEXTURE_WRAP_MIRRORED_REPEAT)
    panel.emissiveTextureTransform[:] = (2.0, 2.0, 0.0, 0.0)
    mask = lambert((1.0, 1.0, 1.0))
    mask.baseColorTextureIndex = 5
    mask.alphaMode = 1  # MASK
    mask.alphaCutoff = 0.5
    blend = lambert((1.0, 0.9, 0.7))
    blend.baseColorTextureIndex = 5
    blend.alphaMode = 2  # BLEND
    blend.baseColorTextureTransform[:] = (2.0, 2.0, 0.0, 0.0)
    s.materials = [floor, back, redm, lambert((0.14, 0.45, 0.09)), lambert((0.73, 0.73, 0.73), emission=15.0),
                   panel, mask, blend, lambert((0.73, 0.73, 0.73))]
    quads = [
        (((0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0)), (0, 0, 1), 0),  # floor
        (((0, 0, 1), (0, 1, 1), (1, 1, 1), (1, 0, 1)), (0, 0, -1), 8),  # ceiling
        (((0, 1, 0), (1, 1, 0), (1, 1, 1), (0, 1, 1)), (0, -1, 0), 1),  # back
        (((0, 0, 0), (0, 1, 0), (0, 1, 1), (0, 0, 1)), (1, 0, 0), 2),  # left (red)
        (((1, 0, 0), (1, 0, 1), (1, 1, 1), (1, 1, 0)), (-1, 0, 0), 3),  # right (green)
    ]
    for corners, normal, mat in quads:
        s.meshes.append(MeshData(*_textured_quad(corners, normal), mat))
    zl = 1.0 - 1e-3
    light = ((0.37, 0.37, zl), (0.63, 0.37, zl), (0.63, 0.63, zl), (0.37, 0.63, zl))
    s.meshes.append(MeshData(*_textured_quad(light, (0, 0, -1)), 4))
    panel_q = ((0.02, 0.55, 0.15), (0.02, 0.85, 0.15), (0.02, 0.85, 0.45), (0.02, 0.55, 0.45))
    s.meshes.append(MeshData(*_textured_quad(panel_q, (1, 0, 0)), 5))
    mask_q = ((0.15, 0.85, 0.05), (0.85, 0.85, 0.05), (0.85, 0.85, 0.75), (0.15, 0.85, 0.05 + 0.7))
    s.meshes.append(MeshData(*_textured_quad(mask_q, (0, -1, 0)), 6))
    zs = 0.82
    sheet = ((0.3, 0.25, zs), (0.7, 0.25, zs), (0.7, 0.65, zs), (0.3, 0.65, zs))
    s.meshes.append(MeshData(*_textured_quad(sheet, (0, 0, -1)), 7))
    s.camera = ((0.5, -1.35, 0.5), (0.5, 0.5, 0.5), (0.0, 0.0, 1.0), 40.0)
    s.environment_color = (0.0, 0.0, 0.0)
    s.environment_strength = 0.0
    return s


# ---------------------------------------------------- differentiable Cornell --
def _subdivide_quad(pos, nrm, n):
    """Splits one quad (corners p0..p3 in order) into n x n sub-quads."""
    p0, p1, p2, p3 = (np.asarray(p, dtype=np.float64) for p in pos)
    out = []
    for i in range(n):
        for j in range(n):
            def at(a, b):
                return (p0 * (1 - a) * (1 - b) + p1 * a * (1 - b) + p2 * a * b + p3 * (1 - a) * b)
            a0, a1, b0, b1 = i / n, (i + 1) / n, j / n, (j + 1) / n
            out.append(_quad(at(a0, b0), at(a1, b0), at(a1, b1), at(a0, b1), nrm[0]))
    return out


PRB_MATERIALS = 64


def differentiable_cornell(seed: int = 0, env_width: int = 512, env_height: int = 256) -> Scene:
    """The Cornell box geometry with one material per face patch for inverse rendering.

    Every wall and box face is split into 2 x 2 patches (17 quads -> 68 patches, one
    mesh each); patch k gets material k mod 63 and the ceiling light material 63, so
    64 materials carry (albedo RGB, roughness, emission RGB). Materials 0..62 are
    coated diffuse (specular 0.5, so roughness matters) with albedo U(0.2, 0.8) and
    roughness U(0.2, 0.8); the light emits 15. The box is lit by the light and by a
    512 x 256 sky (no sun) through its open side. Material 0 of the renderer is its
    lazily created default, so scene material m is renderer material m + 1.
    """
    rng = np.random.default_rng(seed)
    s = Scene()
    for _ in range(PRB_MATERIALS - 1):
        m = default_material(baseColor=tuple(rng.uniform(0.2, 0.8, 3)), roughness=float(rng.uniform(0.2, 0.8)),
                             specular=0.5, metallic=0.0)
        m.emissionLuminance = 0.0
        s.materials.append(m)
    s.materials.append(lambert((0.73, 0.73, 0.73), emission=15.0))
    quads = [
        _quad((0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0), (0, 0, 1)),
        _quad((0, 0, 1), (0, 1, 1), (1, 1, 1), (1, 0, 1), (0, 0, -1)),
        _quad((0, 1, 0), (1, 1, 0), (1, 1, 1), (0, 1, 1), (0, -1, 0)),
        _quad((0, 0, 0), (0, 1, 0), (0, 1, 1), (0, 0, 1), (1, 0, 0)),
        _quad((1, 0, 0), (1, 0, 1), (1, 1, 1), (1, 1, 0), (-1, 0, 0)),
    ] + _box(0.53, 0.83, 0.12, 0.42, 0.3) + _box(0.17, 0.47, 0.48, 0.78, 0.6)
    k = 0
    for pos, nrm, _ in quads:
        for sub in _subdivide_quad(pos, nrm, 2):
            s.meshes.append(MeshData(*_merge([sub]), k % (PRB_MATERIALS - 1)))
            k += 1
    zl = 1.0 - 1e-3
    light = _quad((0.37, 0.37, zl), (0.63, 0.37, zl), (0.63, 0.63, zl), (0.37, 0.63, zl), (0, 0, -1))
    s.meshes.append(MeshData(*_merge([light]), PRB_MATERIALS - 1))
    s.camera = ((0.5, -1.35, 0.5), (0.5, 0.5, 0.5), (0.0, 0.0, 1.0), 40.0)
    s.environment = sky_environment(env_width, env_height, sun_radiance=0.0)
    return s


def prb_views(seed: int = 0, count: int = 3):
    """Seeded camera views into the open side of the box (position, target, up, vfov)."""
    rng = np.random.default_rng(1000 + seed)
    views = []
    for _ in range(count):
        pos = (0.5 + rng.uniform(-0.25, 0.25), -1.35 + rng.uniform(-0.15, 0.15), 0.5 + rng.uniform(-0.2, 0.2))
        target = (0.5 + rng.uniform(-0.1, 0.1), 0.5, 0.5 + rng.uniform(-0.1, 0.1))
        views.append((pos, target, (0.0, 0.0, 1.0), 40.0))
    return views


# ---------------------------------------------------------------- sphere --
_ICO_T = (1.0 + 5.0 ** 0.5) / 2.0
_ICO_V = np.array(
    [[-1, _ICO_T, 0], [1, _ICO_T, 0], [-1, -_ICO_T, 0], [1, -_ICO_T, 0],
     [0, -1, _ICO_T], [0, 1, _ICO_T], [0, -1, -_ICO_T], [0, 1, -_ICO_T],
     [_ICO_T, 0, -1], [_ICO_T, 0, 1], [-_ICO_T, 0, -1], [-_ICO_T, 0, 1]], dtype=np.float64)
_ICO_F = np.array(
    [[0, 11, 5], [0, 5, 1], [0, 1, 7], [0, 7, 10], [0, 10, 11], [1, 5, 9], [5, 11, 4], [11, 10, 2], [10, 7, 6],
     [7, 1, 8], [3, 9, 4], [3, 4, 2], [3, 2, 6], [3, 6, 8], [3, 8, 9], [4, 9, 5], [2, 4, 11], [6, 2, 10], [8, 6, 7],
     [9, 8, 1]], dtype=np.int64)


def geodesic_sphere(freq: int):
    """Frequency-f geodesic subdivision: 20 f^2 triangles, 10 f^2 + 2 unit vertices."""
    f = int(freq)
    # lattice points (i, j) with i + j <= f on each face; vertex = A + (B-A) i/f + (C-A) j/f
    ii, jj = np.meshgrid(np.arange(f + 1), np.arange(f + 1), indexing="ij")
    keep = ii + jj <= f
    ii, jj = ii[keep], jj[keep]
    local = -np.ones((f + 1, f + 1), dtype=np.int64)
    local[ii, jj] = np.arange(len(ii))
    # triangles in the lattice
    a = local[:-1, :-1]
    up_mask = (np.add.outer(np.arange(f), np.arange(f)) <= f - 1)
    i0, j0 = np.nonzero(up_mask)
    tri_up = np.stack([local[i0, j0], local[i0 + 1, j0], local[i0, j0 + 1]], axis=1)
    dn_mask = (np.add.outer(np.arange(f), np.arange(f)) <= f - 2)
    i1, j1 = np.nonzero(dn_mask)
    tri_dn = np.stack([local[i1 + 1, j1], local[i1 + 1, j1 + 1], local[i1, j1 + 1]], axis=1)
    lat_tris = np.concatenate([tri_up, tri_dn])
    del a
    pts, tris = [], []
    base = 0
    for face in _ICO_F:
        A, B, Cc = _ICO_V[face]
        p = A[None] + np.outer(ii / f, B - A) + np.outer(jj / f, Cc - A)
        pts.append(p)
        tris.append(lat_tris + base)
        base += len(p)
    pts = np.concatenate(pts)
    tris = np.concatenate(tris)
    pts /= np.linalg.norm(pts, axis=1, keepdims=True)
    # weld duplicates on shared edges
    key = np.round(pts * (1 << 20)).astype(np.int64)
    _, first, inverse = np.unique(key, axis=0, return_index=True, return_inverse=True)
    verts = pts[first]
    tris = inverse.reshape(-1)[tris]
    return verts, tris.astype(np.int64)


def _value_noise(p: np.ndarray, seed: int, octaves: int = 3, base_freq: float = 2.0) -> np.ndarray:
    """Seeded 3D value noise: lattice values hashed from integer coordinates,
    smoothstep-weighted trilinear interpolation, octaves at doubling frequency
    and halving amplitude; output roughly in [-1, 1]."""
    total = np.zeros(len(p))
    amp, freq, norm = 1.0, base_freq, 0.0
    for o in range(octaves):
        q = p * freq + 17.0 * o
        i0 = np.floor(q).astype(np.int64)
        t = q - i0
        w = t * t * (3.0 - 2.0 * t)

        def lattice(ix, iy, iz):
            h = (ix * 73856093) ^ (iy * 19349663) ^ (iz * 83492791) ^ (seed * 2654435761 + o * 40503)
            h = (h ^ (h >> 13)) * 0x5BD1E995
            h = (h ^ (h >> 15)) & 0xFFFFFF
            return h / float(0xFFFFFF) * 2.0 - 1.0

        acc = 0.0
        for dx in (0, 1):
            for dy in (0, 1):
                for dz in (0, 1):
                    wx = w[:, 0] if dx else 1.0 - w[:, 0]
                    wy = w[:, 1] if dy else 1.0 - w[:, 1]
                    wz = w[:, 2] if dz else 1.0 - w[:, 2]
                    acc = acc + wx * wy * wz * lattice(i0[:, 0] + dx, i0[:, 1] + dy, i0[:, 2] + dz)
        total += amp * acc
        norm += amp
        amp *= 0.5
        freq *= 2.0
    return total / norm


def _vertex_normals(v: np.ndarray, f: np.ndarray) -> np.ndarray:
    fn = np.cross(v[f[:, 1]] - v[f[:, 0]], v[f[:, 2]] - v[f[:, 0]])
    n = np.zeros_like(v)
    for k in range(3):
        np.add.at(n, f[:, k], fn)
    return n / np.maximum(np.linalg.norm(n, axis=1, keepdims=True), 1e-30)


def _face_hash(n: int, seed: int) -> np.ndarray:
    x = (np.arange(n, dtype=np.uint64) * np.uint64(0x9E3779B97F4A7C15)) ^ np.uint64((seed * 0x632BE59BD9B4E019) & 0xFFFFFFFFFFFFFFFF)
    x ^= x >> np.uint64(31)
    x *= np.uint64(0xBF58476D1CE4E5B9)
    x ^= x >> np.uint64(27)
    return x


def sphere_materials(count: int = 64, seed: int = 7, dielectric_fraction: float = 0.2):
    rng = np.random.default_rng(seed)
    mats = []
    kinds = rng.random(count) < dielectric_fraction
    for k in range(count):
        rough = float(rng.uniform(0.05, 1.0))
        albedo = tuple(float(x) for x in rng.uniform(0.05, 0.95, 3))
        if kinds[k]:
            m = default_material(baseColor=albedo, roughness=rough, metallic=0.0, specular=0.5)
            m.transmission = 1.0
            m.ior = 1.5
            m.abbeNumber = 0.0
        else:
            m = default_material(baseColor=albedo, roughness=rough, metallic=0.0, specular=0.5)
        mats.append(m)
    return mats


def displaced_sphere(freq: int = 229, radius: float = 1.0, seed: int = 11, materials: int = 64) -> Scene:
    verts, tris = geodesic_sphere(freq)
    disp = _value_noise(verts * radius, seed) * (0.1 * radius)
    pos = verts * (radius + disp)[:, None]
    nrm = _vertex_normals(pos, tris)
    mat_id = (_face_hash(len(tris), seed) % np.uint64(materials)).astype(np.int64)
    s = Scene()
    s.materials = sphere_materials(materials, seed=seed)
    order = np.argsort(mat_id, kind="stable")
    bounds = np.searchsorted(mat_id[order], np.arange(materials + 1))
    for m in range(materials):
        ft = tris[order[bounds[m]:bounds[m + 1]]]
        if len(ft) == 0:
            continue
        used, local = np.unique(ft, return_inverse=True)
        s.meshes.append(MeshData(make_vertices(pos[used], nrm[used]), local.reshape(-1, 3).astype(np.uint32), m))
    s.camera = ((2.6, -2.2, 1.1), (0.0, 0.0, 0.0), (0.0, 0.0, 1.0), 40.0)
    return s


def sky_environment(width: int = 2048, height: int = 1024, sun_elevation_deg: float = 35.0,
                    sun_azimuth_deg: float = 40.0, sun_radiance: float = 1.0e4,
                    sun_diameter_deg: float = 0.5) -> np.ndarray:
    """Lat-long (z up) sky: zenith-to-horizon gradient, dim ground, sun disc.
    Texel (x, y) covers u = (x + 0.5)/W, v = (y + 0.5)/H with the reference's
    mapping u = frac(atan2(d.y, d.x)/(2 pi) + 0.5), v = acos(d.z)/pi."""
    u = (np.arange(width) + 0.5) / width
    v = (np.arange(height) + 0.5) / height
    phi = (u - 0.5) * 2.0 * np.pi
    theta = v * np.pi
    st, ct = np.sin(theta)[:, None], np.cos(theta)[:, None]
    d = np.stack([st * np.cos(phi)[None, :], st * np.sin(phi)[None, :], np.broadcast_to(ct, (height, width))], axis=-1)
    elev = np.clip(d[..., 2], -1.0, 1.0)
    zenith = np.array([0.15, 0.35, 0.85])
    horizon = np.array([0.75, 0.85, 1.0])
    ground = np.array([0.08, 0.07, 0.06])
    t = np.clip(elev, 0.0, 1.0)[..., None] ** 0.5
    sky = horizon * (1.0 - t) + zenith * t
    rgb = np.where(elev[..., None] >= 0.0, sky, ground)
    se, sa = np.radians(sun_elevation_deg), np.radians(sun_azimuth_deg)
    sun = np.array([np.cos(se) * np.cos(sa), np.cos(se) * np.sin(sa), np.sin(se)])
    cosang = d @ sun
    inside = cosang >= np.cos(np.radians(sun_diameter_deg * 0.5))
    rgb[inside] = sun_radiance
    out = np.ones((height, width, 4), dtype=np.float32)
    out[..., :3] = rgb
    return out


def load(renderer: Renderer, scene: Scene):
    """Creates the scene's materials and meshes through the C API."""
    for i, (pixels, fmt, space) in enumerate(scene.textures):
        renderer.add_texture(pixels, fmt, space, f"t{i}")
    mat_ids = [renderer.add_material(m, f"m{i}") for i, m in enumerate(scene.materials)]
    for mesh in scene.meshes:
        idx = renderer.add_mesh(mesh.vertices, mesh.indices)
        renderer.set_mesh_material(idx, mat_ids[mesh.material])
        if mesh.transform is not None:
            m = np.asarray(mesh.transform, dtype=np.float32)
            renderer.set_mesh_transform_matrix(idx, np.ascontiguousarray(m.T).reshape(-1))
        if mesh.opacity != 1.0:
            renderer.set_mesh_opacity(idx, mesh.opacity)
    pos, target, up, vfov = scene.camera
    renderer.set_camera(pos, target, up, vfov)
    if scene.environment is not None:
        tex = renderer.add_texture_rgba32f(scene.environment)
        renderer.set_environment_texture(tex)
        renderer.set_environment_light((1.0, 1.0, 1.0), 1.0)
    else:
        renderer.set_environment_light(scene.environment_color, scene.environment_strength)


# ---------------------------------------------------------------- stress --
def _random_rotation(rng: np.random.Generator) -> np.ndarray:
    """Uniform random rotation from a normalised Gaussian quaternion."""
    q = rng.normal(size=4)
    w, x, y, z = q / np.linalg.norm(q)
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                     [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                     [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])


def random_triangles(count: int, seed: int = 5, edge=(1e-3, 1e-2)) -> MeshData:
    """`count` independent triangles in the unit cube: a random centre, a random
    orientation and edges of length U(edge) along two random directions."""
    rng = np.random.default_rng(seed)
    c = rng.random((count, 3))
    e1 = rng.normal(size=(count, 3))
    e1 /= np.linalg.norm(e1, axis=1, keepdims=True)
    e2 = rng.normal(size=(count, 3))
    e2 -= np.sum(e2 * e1, axis=1, keepdims=True) * e1
    e2 /= np.linalg.norm(e2, axis=1, keepdims=True)
    l1 = rng.uniform(edge[0], edge[1], (count, 1))
    l2 = rng.uniform(edge[0], edge[1], (count, 1))
    p0 = c - (l1 * e1 + l2 * e2) / 3.0
    pos = np.stack([p0, p0 + l1 * e1, p0 + l2 * e2], axis=1).reshape(-1, 3)
    nrm = np.repeat(np.cross(e1, e2), 3, axis=0)
    verts = make_vertices(pos.astype(np.float32), nrm.astype(np.float32))
    return MeshData(verts, np.arange(3 * count, dtype=np.uint32).reshape(-1, 3), 0)


def stress_scene(instances: int = 16, freq: int = 229, random_count: int = 4_000_000, seed: int = 21) -> Scene:
    """Incoherent-traversal stress configuration: `instances` copies of the
    displaced sphere (one diffuse material, seeded rigid transforms within a
    6 x 6 x 6 region around the unit cube) and `random_count` random small
    triangles in the unit cube; the camera sits among the instances so every
    primary ray hits geometry. freq 229 and 16 instances give 16,781,120 +
    4,000,000 triangles."""
    rng = np.random.default_rng(seed)
    verts, tris = geodesic_sphere(freq)
    disp = _value_noise(verts, 11) * 0.1
    pos = verts * (1.0 + disp)[:, None]
    nrm = _vertex_normals(pos, tris)
    sphere = make_vertices(pos, nrm)
    idx = tris.astype(np.uint32)
    s = Scene()
    s.materials = [lambert((0.7, 0.7, 0.7)), lambert((0.5, 0.6, 0.8))]
    for k in range(instances):
        m = np.eye(4)
        m[:3, :3] = _random_rotation(rng)
        m[:3, 3] = rng.uniform([-2.5, -2.5, -2.5], [3.5, 3.5, 3.5])
        s.meshes.append(MeshData(sphere, idx, 0, m))
    if random_count:
        s.meshes.append(random_triangles(random_count, seed=seed + 1))
        s.meshes[-1].material = 1
    s.camera = ((0.5, -4.5, 1.2), (0.5, 0.5, 0.5), (0.0, 0.0, 1.0), 60.0)
    return s
