"""ORACLE -- test infrastructure only.

ctypes wrapper of liboracle.so, the CPU restatement of the reference's path
tracer (see oracle_core.c). Only tests/, __graft_entry__.smoke() and the
benchmark's CPU-baseline leg use this module, as the checker.
"""

from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")
_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB):
        build()
    L = C.CDLL(LIB)
    P = C.c_void_p
    f32p = C.POINTER(C.c_float)
    f64p = C.POINTER(C.c_double)
    u32p = C.POINTER(C.c_uint32)
    u16p = C.POINTER(C.c_uint16)
    u64p = C.POINTER(C.c_uint64)
    L.oracle_bvh_build.restype = P
    L.oracle_bvh_build.argtypes = [f32p, C.c_uint32]
    L.oracle_bvh_free.argtypes = [P]
    L.oracle_bvh_node_count.restype = C.c_uint32
    L.oracle_bvh_node_count.argtypes = [P]
    L.oracle_build_alias.restype = C.c_int
    L.oracle_build_alias.argtypes = [f32p, C.c_uint32, f32p, u32p]
    for suffix, rp, rt in (("_f32", f32p, C.c_float), ("_f64", f64p, C.c_double)):
        g = lambda n: getattr(L, n + suffix)  # noqa: E731
        g("oracle_render").restype = C.c_int
        g("oracle_render").argtypes = [P, P, C.c_uint32, C.c_uint32, rp, u16p, C.c_uint32, C.c_uint32, C.c_int, u64p]
        g("oracle_render_ex").restype = C.c_int
        g("oracle_render_ex").argtypes = [P, P, C.c_uint32, C.c_uint32, rp, rp, rp, u16p, C.c_uint32, C.c_uint32,
                                          C.c_int, u64p]
        g("oracle_primary_hits").restype = C.c_int
        g("oracle_primary_hits").argtypes = [P, P, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, u32p, u32p, C.c_int]
        g("oracle_trace_rays").restype = C.c_int
        g("oracle_trace_rays").argtypes = [P, P, f32p, C.c_uint64, u32p, u32p, f32p, f32p, C.c_int]
        g("oracle_rand_sequence").argtypes = [C.c_uint32, C.c_uint64, C.c_int, C.c_int, C.c_uint32, C.c_uint32,
                                              C.c_uint32, rp]
        g("oracle_pcg32_raw").argtypes = [C.c_uint64, C.c_uint64, C.c_uint32, u32p]
        g("oracle_hash").restype = C.c_uint32
        g("oracle_hash").argtypes = [C.c_uint32]
        g("oracle_hash_states").argtypes = [C.c_int, C.c_int, C.c_uint32, C.c_uint32, C.c_uint32, u32p]
        g("oracle_sheen_ltc").restype = rt
        g("oracle_sheen_ltc").argtypes = [rt, rt, C.c_uint32]
        g("oracle_sheen_eval").argtypes = [rt, rp, rp, rp]
        g("oracle_camera_ray").argtypes = [P, C.c_int, C.c_int, rt, rt, rp]
        g("oracle_intersect").restype = C.c_int
        g("oracle_intersect").argtypes = [f32p, rp, rp]
        g("oracle_bsdf_eval").argtypes = [P, P, rp, rp, C.c_uint32, rt, C.c_uint32, rp, rp]
        g("oracle_sample_material_slot").argtypes = [P, P, C.c_uint32, f32p, C.c_uint32, f32p]
        g("oracle_bsdf_eval_hero").argtypes = [P, P, rp, rp, C.c_uint32, rp, rp, rp]
        g("oracle_bsdf_sample_batch").argtypes = [P, P, rp, C.c_uint32, rt, C.c_uint32, C.c_uint32, C.c_uint32, rp]
        g("oracle_rgb2spec").restype = rt
        g("oracle_rgb2spec").argtypes = [P, rp, rt]
        g("oracle_cie_xyz").argtypes = [rt, rp]
        g("oracle_xyz_to_srgb").argtypes = [rp, rp]
        g("oracle_dispersive_ior").restype = rt
        g("oracle_dispersive_ior").argtypes = [rt, rt, rt]
        g("oracle_fresnel_dielectric").restype = rt
        g("oracle_fresnel_dielectric").argtypes = [rt, rt]
        g("oracle_environment").argtypes = [P, rp, rp]
        g("oracle_hero_wavelengths").argtypes = [C.c_uint32, C.c_uint32, rp]
        g("oracle_env_sample_batch").argtypes = [P, C.c_uint64, C.c_uint32, rp]
        g("oracle_env_pdf").restype = rt
        g("oracle_env_pdf").argtypes = [P, rp]
    _lib = L
    return L


def _dtype(precision):
    return (np.float32, "_f32") if precision == 32 else (np.float64, "_f64")


def ptr(a):
    if a.dtype == np.float32:
        return a.ctypes.data_as(C.POINTER(C.c_float))
    if a.dtype == np.float64:
        return a.ctypes.data_as(C.POINTER(C.c_double))
    if a.dtype == np.uint32:
        return a.ctypes.data_as(C.POINTER(C.c_uint32))
    if a.dtype == np.uint16:
        return a.ctypes.data_as(C.POINTER(C.c_uint16))
    raise TypeError(a.dtype)


class OracleScene:
    """Holds exported scene buffers (kept alive by the renderer) and an oracle BVH."""

    def __init__(self, buffers, use_bvh: bool = True, **scene_overrides):
        """Snapshots the SceneData uniform (the renderer advances frameNumber
        after every frame); keyword arguments override its fields."""
        import vkrt_amd

        self.buffers = buffers
        self.scene = vkrt_amd.SceneData()
        C.memmove(C.addressof(self.scene), buffers.sceneData, C.sizeof(vkrt_amd.SceneData))
        for key, value in scene_overrides.items():
            setattr(self.scene, key, value)
        self.buffers.sceneData = C.pointer(self.scene)
        self.addr = C.addressof(buffers)
        self.bvh = None
        n = buffers.triangleCount
        if use_bvh and n > 0:
            self.bvh = lib().oracle_bvh_build(buffers.worldTriangles, n)

    def close(self):
        if self.bvh:
            lib().oracle_bvh_free(self.bvh)
            self.bvh = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def render(self, width, height, accum=None, precision=32, rows=None, threads=None, output=False, features=None):
        """One frame into accum (H x W x 4, .w = samples). features, if given, is a
        pair of H x W x 4 albedo and normal images (.w = feature weight) that the
        frame blends into, as the reference's albedo / normal outputs."""
        dt, sfx = _dtype(precision)
        if accum is None:
            accum = np.zeros((height, width, 4), dtype=dt)
        assert accum.dtype == dt and accum.shape == (height, width, 4)
        albedo, normal = features if features is not None else (None, None)
        for img in (albedo, normal):
            assert img is None or (img.dtype == dt and img.shape == (height, width, 4))
        out = np.zeros((height, width, 4), dtype=np.uint16) if output else None
        r0, r1 = rows if rows is not None else (0, height)
        bounces = C.c_uint64()
        rc = getattr(lib(), "oracle_render_ex" + sfx)(
            self.addr, self.bvh, width, height, ptr(accum), ptr(albedo) if albedo is not None else None,
            ptr(normal) if normal is not None else None, ptr(out) if out is not None else None, r0, r1,
            threads or default_threads(), C.byref(bounces))
        assert rc == 0
        return (accum, out, bounces.value) if output else (accum, bounces.value)

    def primary_hits(self, width, height, sample_index=0, prev_samples=0, precision=32, threads=None):
        mesh = np.zeros(width * height, dtype=np.uint32)
        prim = np.zeros(width * height, dtype=np.uint32)
        _, sfx = _dtype(precision)
        rc = getattr(lib(), "oracle_primary_hits" + sfx)(self.addr, self.bvh, width, height, sample_index,
                                                         prev_samples, ptr(mesh), ptr(prim), threads or default_threads())
        assert rc == 0
        return mesh, prim

    def trace_rays(self, rays, precision=32, threads=None):
        rays = np.ascontiguousarray(rays, dtype=np.float32)
        n = len(rays)
        mesh = np.zeros(n, dtype=np.uint32)
        prim = np.zeros(n, dtype=np.uint32)
        t = np.zeros(n, dtype=np.float32)
        bary = np.zeros((n, 2), dtype=np.float32)
        _, sfx = _dtype(precision)
        getattr(lib(), "oracle_trace_rays" + sfx)(self.addr, self.bvh, ptr(rays), n, ptr(mesh), ptr(prim), ptr(t),
                                                  ptr(bary), threads or default_threads())
        return mesh, prim, t, bary


def default_threads():
    """OpenMP threads for the oracle: the cores this process may use, capped by
    OMP_NUM_THREADS when the host share is set."""
    n = len(os.sched_getaffinity(0))
    share = int(os.environ.get("OMP_NUM_THREADS", "0") or 0)
    return max(1, min(n, share) if share > 0 else n)


def sheen_ltc(cos_theta, roughness, layer):
    return float(lib().oracle_sheen_ltc_f64(cos_theta, roughness, layer))


def sheen_eval(roughness, wo, wi):
    """The bare sheen lobe, unit colour: (albedo * density, density)."""
    out = np.zeros(2)
    lib().oracle_sheen_eval_f64(roughness, ptr(np.ascontiguousarray(wo, dtype=np.float64)),
                                ptr(np.ascontiguousarray(wi, dtype=np.float64)), ptr(out))
    return float(out[0]), float(out[1])


def rand_sequence(policy, seed, px, py, frame, sample, n, precision=32):
    dt, sfx = _dtype(precision)
    out = np.zeros(n, dtype=dt)
    getattr(lib(), "oracle_rand_sequence" + sfx)(policy, seed, px, py, frame, sample, n, ptr(out))
    return out


def pcg32_raw(init_state, init_seq, n):
    out = np.zeros(n, dtype=np.uint32)
    lib().oracle_pcg32_raw_f32(C.c_uint64(init_state), C.c_uint64(init_seq), n, ptr(out))
    return out


def oracle_hash(v):
    return lib().oracle_hash_f32(v)


def hash_states(px, py, frame, sample, n):
    """Hash-policy RNG states: the pixel seed, then the state after each draw."""
    out = np.zeros(n, dtype=np.uint32)
    lib().oracle_hash_states_f32(px, py, frame, sample, n, ptr(out))
    return out


def intersect(tri9, ray8, precision=32):
    dt, sfx = _dtype(precision)
    tri = np.ascontiguousarray(tri9, dtype=np.float32)
    ray = np.ascontiguousarray(ray8, dtype=dt)
    out = np.zeros(3, dtype=dt)
    hit = getattr(lib(), "oracle_intersect" + sfx)(ptr(tri), ptr(ray), ptr(out))
    return bool(hit), out


def bsdf_eval(buffers, material, wo, wi, front_face=1, wavelength=0.0, mode=0, precision=64):
    dt, sfx = _dtype(precision)
    wo = np.ascontiguousarray(wo, dtype=dt)
    wi = np.ascontiguousarray(wi, dtype=dt)
    val = np.zeros(3, dtype=dt)
    pdf = np.zeros(1, dtype=dt)
    getattr(lib(), "oracle_bsdf_eval" + sfx)(C.addressof(buffers) if buffers is not None else None,
                                            C.addressof(material), ptr(wo), ptr(wi), front_face, wavelength, mode,
                                            ptr(val), ptr(pdf))
    return val, float(pdf[0])


def bsdf_eval_hero(buffers, material, wo, wi, lambdas, front_face=1, precision=64):
    dt, sfx = _dtype(precision)
    wo = np.ascontiguousarray(wo, dtype=dt)
    wi = np.ascontiguousarray(wi, dtype=dt)
    lam = np.ascontiguousarray(lambdas, dtype=dt)
    val = np.zeros(4, dtype=dt)
    pdf = np.zeros(4, dtype=dt)
    getattr(lib(), "oracle_bsdf_eval_hero" + sfx)(C.addressof(buffers), C.addressof(material), ptr(wo), ptr(wi),
                                                 front_face, ptr(lam), ptr(val), ptr(pdf))
    return val, pdf


def bsdf_sample_batch(buffers, material, wo, n, seed=1, front_face=1, wavelength=0.0, mode=0, precision=64):
    dt, sfx = _dtype(precision)
    wo = np.ascontiguousarray(wo, dtype=dt)
    out = np.zeros((n, 8), dtype=dt)
    getattr(lib(), "oracle_bsdf_sample_batch" + sfx)(C.addressof(buffers) if buffers is not None else None,
                                                    C.addressof(material), ptr(wo), front_face, wavelength, mode, seed,
                                                    n, ptr(out))
    return out


def sample_material_slot(buffers, material, slot, uv):
    """A material texture slot sampled at the N x 4 coordinates uv = (u0, v0,
    u1, v1) (both texture-coordinate sets); returns N x 4 float32."""
    uv = np.ascontiguousarray(uv, dtype=np.float32).reshape(-1, 4)
    out = np.zeros((len(uv), 4), dtype=np.float32)
    lib().oracle_sample_material_slot_f32(C.addressof(buffers), C.addressof(material), slot, ptr(uv), len(uv),
                                          ptr(out))
    return out


def rgb2spec(buffers, rgb, wavelength, precision=64):
    dt, sfx = _dtype(precision)
    rgb = np.ascontiguousarray(rgb, dtype=dt)
    return float(getattr(lib(), "oracle_rgb2spec" + sfx)(C.addressof(buffers), ptr(rgb), wavelength))


def cie_xyz(wavelength, precision=64):
    dt, sfx = _dtype(precision)
    out = np.zeros(3, dtype=dt)
    getattr(lib(), "oracle_cie_xyz" + sfx)(wavelength, ptr(out))
    return out


def xyz_to_srgb(xyz, precision=64):
    dt, sfx = _dtype(precision)
    x = np.ascontiguousarray(xyz, dtype=dt)
    out = np.zeros(3, dtype=dt)
    getattr(lib(), "oracle_xyz_to_srgb" + sfx)(ptr(x), ptr(out))
    return out


def dispersive_ior(ior, abbe, wavelength, precision=64):
    _, sfx = _dtype(precision)
    return float(getattr(lib(), "oracle_dispersive_ior" + sfx)(ior, abbe, wavelength))


def fresnel_dielectric(cos_theta, eta, precision=64):
    _, sfx = _dtype(precision)
    return float(getattr(lib(), "oracle_fresnel_dielectric" + sfx)(cos_theta, eta))


def environment(buffers, direction, precision=64):
    dt, sfx = _dtype(precision)
    d = np.ascontiguousarray(direction, dtype=dt)
    out = np.zeros(3, dtype=dt)
    getattr(lib(), "oracle_environment" + sfx)(C.addressof(buffers), ptr(d), ptr(out))
    return out


def hero_wavelengths(hash_state, sample_index, precision=64):
    dt, sfx = _dtype(precision)
    out = np.zeros(4, dtype=dt)
    getattr(lib(), "oracle_hero_wavelengths" + sfx)(hash_state, sample_index, ptr(out))
    return out


def env_sample_batch(buffers, n, seed=1, precision=64):
    dt, sfx = _dtype(precision)
    out = np.zeros((n, 6), dtype=dt)
    getattr(lib(), "oracle_env_sample_batch" + sfx)(C.addressof(buffers), seed, n, ptr(out))
    return out


def env_pdf(buffers, direction, precision=64):
    dt, sfx = _dtype(precision)
    d = np.ascontiguousarray(direction, dtype=dt)
    return float(getattr(lib(), "oracle_env_pdf" + sfx)(C.addressof(buffers), ptr(d)))


def build_alias(pmf):
    pmf = np.ascontiguousarray(pmf, dtype=np.float32)
    q = np.zeros(len(pmf), dtype=np.float32)
    idx = np.zeros(len(pmf), dtype=np.uint32)
    assert lib().oracle_build_alias(ptr(pmf), len(pmf), ptr(q), ptr(idx)) == 1
    return q, idx


def relmse(x, ref, eps=1e-2):
    """mean((x - ref)^2 / (ref^2 + eps)) over pixels and channels."""
    x = np.asarray(x, dtype=np.float64)
    ref = np.asarray(ref, dtype=np.float64)
    return float(np.mean((x - ref) ** 2 / (ref ** 2 + eps)))
