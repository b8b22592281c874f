// Shading device code for the wavefront path tracer: random streams, camera
// rays, surface reconstruction, rgb2spec spectral uplift, the layered
// principled BSDF (GGX / diffuse / dielectric / clearcoat / fake subsurface),
// emissive-light and environment sampling. Each block names the reference
// shader it reproduces; operation order follows those shaders so results
// agree with them at the float level (compiled with FMA contraction off).
#pragma once

#include "dev_common.h"
#include "dev_texture.h"

namespace mrt {

// ---------------------------------------------------------------- RNG --
// Policy 0: src/shaders/sampling/random.slang:7-27 (integer hash).
// Policy 1: PCG32 stream per path, seeded by a 64-bit seed and selected by
// (pixel, frame, sample). Both draw 24-bit uniforms like random.slang:16-19.
struct Rng {
    uint64_t state;
    uint64_t inc;
    uint32_t policy;
};

DEV uint32_t hash32(uint32_t v) {
    v ^= v >> 16;
    v *= 0x7feb352du;
    v ^= v >> 15;
    v *= 0x846ca68bu;
    v ^= v >> 16;
    return v;
}

DEV uint64_t mix64(uint64_t z) {
    z += 0x9e3779b97f4a7c15ull;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

DEV uint32_t pcgStep(Rng& r) {
    uint64_t old = r.state;
    r.state = old * 6364136223846793005ull + r.inc;
    uint32_t xs = (uint32_t)(((old >> 18u) ^ old) >> 27u);
    uint32_t rot = (uint32_t)(old >> 59u);
    return (xs >> rot) | (xs << ((32u - rot) & 31u));
}

DEV uint64_t pcgStream(int px, int py, uint32_t frame, uint32_t sampleIndex) {
    return mix64((((uint64_t)(uint32_t)px) << 32) | (uint32_t)py) ^ mix64((((uint64_t)frame) << 32) | sampleIndex);
}

DEV Rng rngInit(uint32_t policy, uint64_t seed, int px, int py, uint32_t frame, uint32_t sampleIndex) {
    Rng r;
    r.policy = policy;
    r.inc = 0;
    if (policy == 0u) {
        uint32_t s = (uint32_t)px * 73856093u;
        s ^= (uint32_t)py * 19349663u;
        s ^= frame * 83492791u;
        s ^= sampleIndex * 2654435761u;
        r.state = hash32(s);
    } else {
        r.inc = (pcgStream(px, py, frame, sampleIndex) << 1u) | 1u;
        r.state = 0u;
        pcgStep(r);
        r.state += seed;
        pcgStep(r);
    }
    return r;
}

DEV float rnd(Rng& r) {
    uint32_t bits;
    if (r.policy == 0u) {
        uint32_t h = hash32((uint32_t)r.state + 0x9e3779b9u);
        r.state = h;
        bits = h & 0x00ffffffu;
    } else {
        bits = pcgStep(r) >> 8u;
    }
    return (float)bits * (1.0f / 16777216.0f);
}

// --------------------------------------------------------------- camera --
DEV F4 mulM4(const vkrt_m4 m, F4 v) {
    return F4{m[0][0] * v.x + m[1][0] * v.y + m[2][0] * v.z + m[3][0] * v.w,
              m[0][1] * v.x + m[1][1] * v.y + m[2][1] * v.z + m[3][1] * v.w,
              m[0][2] * v.x + m[1][2] * v.y + m[2][2] * v.z + m[3][2] * v.w,
              m[0][3] * v.x + m[1][3] * v.y + m[2][3] * v.z + m[3][3] * v.w};
}

// src/shaders/camera/ray.slang:15-30
DEV void primaryRay(const SceneData& sd, int px, int py, float jx, float jy, F3& o, F3& d) {
    float vx = (float)(px - (int)sd.viewportRect[0]) + 0.5f + jx;
    float vy = (float)(py - (int)sd.viewportRect[1]) + 0.5f + jy;
    float u = vx / (float)sd.viewportRect[2];
    float v = vy / (float)sd.viewportRect[3];
    float nx = u * 2.0f - 1.0f, ny = v * 2.0f - 1.0f;
    F4 vd = mulM4(sd.projInverse, F4{nx, ny, 1.0f, 1.0f});
    F4 org = mulM4(sd.viewInverse, F4{0.0f, 0.0f, 0.0f, 1.0f});
    F4 dir = mulM4(sd.viewInverse, F4{vd.x, vd.y, vd.z, 0.0f});
    o = F3{org.x, org.y, org.z};
    d = normalize(F3{dir.x, dir.y, dir.z});
}

// src/shaders/camera/ray.slang:32-39
DEV F3 safeNormalize(F3 v) {
    float l2 = dot(v, v);
    if (l2 <= kSafeNormalizeEps) return F3{0.0f, 0.0f, 1.0f};
    return v * rsq(l2);
}

// -------------------------------------------------------------- packing --
// src/shaders/geometry/packing.slang:13-56
DEV F3 decodeOct(float px, float py) {
    float z = 1.0f - fabsf(px) - fabsf(py);
    if (z < 0.0f) {
        float ox = px;
        px = (1.0f - fabsf(py)) * (ox >= 0.0f ? 1.0f : -1.0f);
        py = (1.0f - fabsf(ox)) * (py >= 0.0f ? 1.0f : -1.0f);
    }
    return normalize(F3{px, py, z});
}
DEV F3 unpackOctNormal(uint32_t p) {
    int32_t x = (int32_t)(p << 16) >> 16;
    int32_t y = (int32_t)p >> 16;
    return decodeOct((float)x / 32767.0f, (float)y / 32767.0f);
}
DEV float unpackSnorm15(uint32_t v) {
    int32_t s = (int32_t)(v & 0x7fffu);
    if (s & 0x4000) s |= ~(int32_t)0x7fff;
    return clampf((float)s / 16383.0f, -1.0f, 1.0f);
}
DEV F4 unpackOctTangent(uint32_t p) {
    F3 d = decodeOct(unpackSnorm15(p), unpackSnorm15(p >> 15u));
    return F4{d.x, d.y, d.z, (p & 0x80000000u) ? -1.0f : 1.0f};
}
DEV F4 unpackColor(uint32_t p) {
    const float s = 1.0f / 255.0f;
    return F4{(float)(p & 0xffu) * s, (float)((p >> 8) & 0xffu) * s, (float)((p >> 16) & 0xffu) * s,
              (float)((p >> 24) & 0xffu) * s};
}

// ----------------------------------------------------------- transforms --
// src/shaders/geometry/surface/transform.slang:9-55
DEV F3 rotateMesh(F3 v, const MeshInfo& m) {
    const float d2r = 0.01745329251994329577f;
    float ax = m.rotation[0] * d2r, ay = m.rotation[1] * d2r, az = m.rotation[2] * d2r;
    float s = sinf(ax), c = cosf(ax);
    v = F3{v.x, c * v.y - s * v.z, s * v.y + c * v.z};
    s = sinf(ay);
    c = cosf(ay);
    v = F3{c * v.x + s * v.z, v.y, -s * v.x + c * v.z};
    s = sinf(az);
    c = cosf(az);
    return F3{c * v.x - s * v.y, s * v.x + c * v.y, v.z};
}
DEV float safeSignedRcp(float v) {
    if (fabsf(v) > 1e-6f) return 1.0f / v;
    return v < 0.0f ? -1.0f / 1e-6f : 1.0f / 1e-6f;
}
DEV F3 meshVector(const MeshInfo& m, F3 v) { return rotateMesh(v * F3{m.scale[0], m.scale[1], m.scale[2]}, m); }
DEV F3 meshNormal(const MeshInfo& m, F3 n) {
    F3 inv = F3{safeSignedRcp(m.scale[0]), safeSignedRcp(m.scale[1]), safeSignedRcp(m.scale[2])};
    return safeNormalize(rotateMesh(n * inv, m));
}

// -------------------------------------------------------------- surface --
struct Basis {
    F3 t, b, n;
};
DEV F3 toWorld(F3 l, const Basis& b) { return b.t * l.x + b.b * l.y + b.n * l.z; }
DEV F3 toLocal(F3 w, const Basis& b) { return F3{dot(w, b.t), dot(w, b.b), dot(w, b.n)}; }

// src/shaders/bsdf/base/basis.slang:10-39
DEV Basis makeBasis(F3 normal, F4 tangent) {
    Basis b;
    b.n = safeNormalize(normal);
    F3 td = F3{tangent.x, tangent.y, tangent.z};
    F3 t = td - b.n * dot(td, b.n);
    if (dot(t, t) <= kSafeNormalizeEps) {
        F3 up = fabsf(b.n.z) < 0.999f ? F3{0.0f, 0.0f, 1.0f} : F3{1.0f, 0.0f, 0.0f};
        t = safeNormalize(cross(up, b.n));
    } else {
        t = safeNormalize(t);
    }
    b.t = t;
    b.b = cross(b.n, b.t) * (tangent.w < 0.0f ? -1.0f : 1.0f);
    return b;
}

struct Surface {
    F3 p;
    F3 ng;  // geometric normal facing the incoming ray
    Basis basis;
    uint32_t frontFace;
    uint32_t mesh;
    uint32_t materialIndex;
    F3 vertexColor;
    bool textured;  // the material samples textures (TEX reconstruction only)
    TexCoords tc;
};

DEV F3 bary3(F3 a, F3 b, F3 c, float u, float v) {
    float w = 1.0f - u - v;
    return a * w + b * u + c * v;
}
DEV F4 bary4(F4 a, F4 b, F4 c, float u, float v) {
    float w = 1.0f - u - v;
    return a * w + b * u + c * v;
}
DEV F4 ldVertexPos(const ShaderVertex& v) { return F4{v.position[0], v.position[1], v.position[2], v.position[3]}; }

// src/shaders/integrator/path/surface_state.slang:14-30 with
// src/shaders/geometry/surface/reconstruct.slang:8-49 (no normal texture)
// The per-triangle part of surface reconstruction (src/shaders/geometry/surface/
// reconstruct.slang:8-49): mesh and the world geometric normal from the object
// edges through the mesh transform, plus the packed vertex attributes.
DEV ShadeTri buildShadeTri(const SceneView& s, uint32_t tri) {
    ShadeTri r;
    uint32_t mesh = s.triMesh[tri];
    const MeshInfo m = s.meshes[mesh];
    uint32_t prim = tri - s.meshTriBase[mesh];
    uint32_t tb = m.indexBase + prim * 3u;
    const ShaderVertex v0 = s.vertices[s.indices[tb + 0] + m.vertexBase];
    const ShaderVertex v1 = s.vertices[s.indices[tb + 1] + m.vertexBase];
    const ShaderVertex v2 = s.vertices[s.indices[tb + 2] + m.vertexBase];
    F4 p0 = ldVertexPos(v0), p1 = ldVertexPos(v1), p2 = ldVertexPos(v2);
    F3 e1 = meshVector(m, F3{p1.x - p0.x, p1.y - p0.y, p1.z - p0.z});
    F3 e2 = meshVector(m, F3{p2.x - p0.x, p2.y - p0.y, p2.z - p0.z});
    F3 gn = safeNormalize(cross(e1, e2));
    if (m.scale[0] * m.scale[1] * m.scale[2] < 0.0f) gn = -gn;
    r.mesh = mesh;
    r.gn[0] = gn.x;
    r.gn[1] = gn.y;
    r.gn[2] = gn.z;
    r.normal[0] = v0.packedNormal;
    r.normal[1] = v1.packedNormal;
    r.normal[2] = v2.packedNormal;
    r.tangent[0] = v0.packedTangent;
    r.tangent[1] = v1.packedTangent;
    r.tangent[2] = v2.packedTangent;
    r.color[0] = v0.packedColor;
    r.color[1] = v1.packedColor;
    r.color[2] = v2.packedColor;
    r.pad[0] = r.pad[1] = r.pad[2] = 0u;
    return r;
}

// TEX: the material's normal texture perturbs the shading normal before it is
// sanitised (surface_state.slang:24-27, textures.slang:143-155), and the
// texture coordinates are kept for the material's texture stage.
template <bool TEX = false>
DEV Surface reconstruct(const SceneView& s, F3 o, F3 d, const Hit& h) {
    Surface sf;
    const uint4* w = reinterpret_cast<const uint4*>(s.shadeTris + h.tri);
    const uint4 w0 = w[0], w1 = w[1], w2 = w[2], w3 = w[3];
    const uint32_t mesh = w0.x;
    const F3 gn = F3{__uint_as_float(w0.y), __uint_as_float(w0.z), __uint_as_float(w0.w)};
    // normals w1.xyz, tangents w1.w w2.xy, colours w2.zw w3.x
    const MeshInfo m = s.meshes[mesh];
    sf.p = o + d * h.t;
    F3 objN = safeNormalize(bary3(unpackOctNormal(w1.x), unpackOctNormal(w1.y), unpackOctNormal(w1.z), h.u, h.v));
    F4 objT = bary4(unpackOctTangent(w1.w), unpackOctTangent(w2.x), unpackOctTangent(w2.y), h.u, h.v);
    float hand = objT.w < 0.0f ? -1.0f : 1.0f;
    F3 snU = meshNormal(m, objN);
    float facing = dot(snU, d) > 0.0f ? -1.0f : 1.0f;
    sf.mesh = mesh;
    sf.materialIndex = m.materialIndex;
    sf.frontFace = dot(gn, d) < 0.0f ? 1u : 0u;
    F3 sn = snU * facing;
    sf.ng = sf.frontFace ? gn : -gn;
    F3 tw = safeNormalize(meshVector(m, F3{objT.x, objT.y, objT.z})) * facing;
    F4 tangent = F4{tw.x, tw.y, tw.z, hand};
    F4 col = bary4(unpackColor(w2.z), unpackColor(w2.w), unpackColor(w3.x), h.u, h.v);
    sf.vertexColor = F3{col.x, col.y, col.z};
    Basis unperturbed = makeBasis(sn, tangent);
    sn = unperturbed.n;
    sf.textured = false;
    if (TEX && s.texels) {
        const Material& mt = s.materials[m.materialIndex];
        if (materialHasTextures(mt)) {
            sf.textured = true;
            sf.tc = hitTexCoords(s, h.tri, mesh, h.u, h.v);
            if (mt.normalTextureIndex != VKRT_INVALID_INDEX) {
                const F4 ns = sampleMaterialSlot(s, mt, VKRT_MATERIAL_TEXTURE_SLOT_NORMAL, sf.tc, f4(0.0f));
                F3 t = F3{ns.x * 2.0f - 1.0f, ns.y * 2.0f - 1.0f, ns.z * 2.0f - 1.0f};
                t.x *= mt.normalTextureScale;
                t.y *= mt.normalTextureScale;
                t = safeNormalize(t);
                sn = safeNormalize(unperturbed.t * t.x + unperturbed.b * t.y + unperturbed.n * t.z);
            }
        }
    }
    {  // sanitizeShadingNormal (basis.slang:29-39)
        F3 n = safeNormalize(sn), g = safeNormalize(sf.ng), out = safeNormalize(-d);
        sn = (dot(n, g) <= 1e-4f || dot(n, out) <= 1e-4f) ? g : n;
    }
    sf.basis = makeBasis(sn, tangent);
    return sf;
}

// ------------------------------------------------------------- spectral --
DEV float lum(F3 c) { return c.x * 0.2126f + c.y * 0.7152f + c.z * 0.0722f; }

struct Spec {
    const float* table;
    uint32_t res, scaleOffset, dataOffset;
};

// src/shaders/utility/rgb2spec.slang:14-90. The table resolution is fixed at
// compile time (the shipped table is 64^3; the host rejects others): constant
// strides let the 24 corner loads share two base addresses, which keeps the
// hero shading kernel within its register budget.
constexpr uint32_t kSpecRes = 64u;
static_assert(kSpecRes == 64u, "specInterval searches 8 x 8 scale entries");
DEV uint32_t specInterval(const Spec& sp, float x) {
    // The largest i in [0, res - 2] with scale[i] <= x, which the reference finds
    // by binary search (rgb2spec.slang:14-30): six dependent loads. The scale is
    // increasing, so two levels of eight-way comparisons give the same index
    // with two dependent steps; the first level's addresses are the same for
    // every lane.
    const float* sc = sp.table + sp.scaleOffset;
    uint32_t b = 0;
#pragma unroll
    for (uint32_t k = 1; k < 8u; k++) b += sc[8u * k] <= x ? 1u : 0u;
    const float* blk = sc + 8u * b;
    uint32_t i = 8u * b;
#pragma unroll
    for (uint32_t k = 1; k < 8u; k++) i += blk[k] <= x ? 1u : 0u;
    return min(i, kSpecRes - 2u);
}

DEV F3 specFetch(const Spec& sp, F3 rgb) {
    float z = fmaxf(rgb.x, fmaxf(rgb.y, rgb.z));
    if (z <= 1e-8f) return f3(0.0f);
    float ch[3] = {rgb.x, rgb.y, rgb.z};
    uint32_t dom = 0;
    if (ch[1] >= ch[dom]) dom = 1;
    if (ch[2] >= ch[dom]) dom = 2;
    float xyScale = (float)(kSpecRes - 1u) / z;
    float x = ch[dom == 2 ? 0 : dom + 1] * xyScale;
    float y = ch[dom == 0 ? 2 : dom - 1] * xyScale;
    uint32_t xi = min((uint32_t)x, kSpecRes - 2u);
    uint32_t yi = min((uint32_t)y, kSpecRes - 2u);
    uint32_t zi = specInterval(sp, z);
    uint32_t off = (((dom * kSpecRes + zi) * kSpecRes + yi) * kSpecRes + xi) * 3u;
    uint32_t dx = 3u, dy = 3u * kSpecRes, dz = dy * kSpecRes;
    float x1 = x - (float)xi, y1 = y - (float)yi, x0 = 1.0f - x1, y0 = 1.0f - y1;
    float s0 = sp.table[sp.scaleOffset + zi], s1 = sp.table[sp.scaleOffset + zi + 1u];
    float z1 = (z - s0) / fmaxf(s1 - s0, 1e-8f);
    float z0 = 1.0f - z1;
    float co[3];
#pragma unroll
    for (int j = 0; j < 3; j++) {
        const float* t = sp.table + sp.dataOffset + off + j;
        co[j] = ((t[0] * x0 + t[dx] * x1) * y0 + (t[dy] * x0 + t[dy + dx] * x1) * y1) * z0 +
                ((t[dz] * x0 + t[dz + dx] * x1) * y0 + (t[dz + dy] * x0 + t[dz + dy + dx] * x1) * y1) * z1;
    }
    return F3{co[0], co[1], co[2]};
}
DEV float specEval(F3 c, float l) {
    float x = (c.x * l + c.y) * l + c.z;
    return 0.5f * x * rsq(x * x + 1.0f) + 0.5f;
}
// src/shaders/utility/spectral.slang:37-64
DEV float spec1(const Spec& sp, F3 rgb, float l) {
    float mx = fmaxf(rgb.x, fmaxf(rgb.y, rgb.z));
    if (mx <= 0.0f) return 0.0f;
    if (mx <= 1.0f) return specEval(specFetch(sp, rgb), l);
    return mx * specEval(specFetch(sp, rgb / mx), l);
}
DEV F4 spec4(const Spec& sp, F3 rgb, F4 l) {
    float mx = fmaxf(rgb.x, fmaxf(rgb.y, rgb.z));
    if (mx <= 0.0f) return f4(0.0f);
    float scale = mx <= 1.0f ? 1.0f : mx;
    F3 c = specFetch(sp, rgb / scale);
    return F4{specEval(c, l.x), specEval(c, l.y), specEval(c, l.z), specEval(c, l.w)} * scale;
}

// src/shaders/utility/spectral.slang:66-118
// exp(-t²/2) = exp2(-(t·k)²) with k = sqrt(log2(e)/2) folded into the slopes:
// one hardware exp2 per lobe instead of expf's extended-precision range reduction
constexpr float kLobeScale = 0.849321800288019f;
DEV float lobe(float l, float mu, float a, float b) {
    float t = (l - mu) * (l < mu ? a * kLobeScale : b * kLobeScale);
    return exp2f(-(t * t));
}
DEV F3 cie(float l) {
    float x = 0.362f * lobe(l, 442.0f, 0.0624f, 0.0374f) + 1.056f * lobe(l, 599.8f, 0.0264f, 0.0323f) -
              0.065f * lobe(l, 501.1f, 0.0490f, 0.0382f);
    float y = 0.821f * lobe(l, 568.8f, 0.0213f, 0.0247f) + 0.286f * lobe(l, 530.9f, 0.0613f, 0.0322f);
    float z = 1.217f * lobe(l, 437.0f, 0.0845f, 0.0278f) + 0.681f * lobe(l, 459.0f, 0.0385f, 0.0725f);
    return F3{fmaxf(x, 0.0f), fmaxf(y, 0.0f), fmaxf(z, 0.0f)};
}
constexpr float kCieYIntegral = 106.9461715f;
constexpr float kInvCieYIntegral = 1.0f / kCieYIntegral;
DEV F3 toXYZ1(float v, float l, float invPdf) { return cie(l) * ((v * invPdf) * kInvCieYIntegral); }
DEV F3 toXYZ4(F4 v, F4 l, F4 ip) {
    F3 r = f3(0.0f);
    r = r + cie(l.x) * ((v.x * ip.x) * kInvCieYIntegral);
    r = r + cie(l.y) * ((v.y * ip.y) * kInvCieYIntegral);
    r = r + cie(l.z) * ((v.z * ip.z) * kInvCieYIntegral);
    r = r + cie(l.w) * ((v.w * ip.w) * kInvCieYIntegral);
    return r;
}
DEV F3 xyzToSrgb(F3 c) {
    F3 lms = F3{0.8951f * c.x + 0.2664f * c.y + -0.1614f * c.z, -0.7502f * c.x + 1.7135f * c.y + 0.0367f * c.z,
                0.0389f * c.x + -0.0685f * c.y + 1.0296f * c.z};
    lms = lms * F3{0.9413344f, 1.0404175f, 1.0895327f};
    F3 d = F3{0.9869929f * lms.x + -0.1470543f * lms.y + 0.1599627f * lms.z,
              0.4323053f * lms.x + 0.5183603f * lms.y + 0.0492912f * lms.z,
              -0.0085287f * lms.x + 0.0400428f * lms.y + 0.9684867f * lms.z};
    return F3{3.2404542f * d.x + -1.5371385f * d.y + -0.4985314f * d.z,
              -0.9692660f * d.x + 1.8760108f * d.y + 0.0415560f * d.z,
              0.0556434f * d.x + -0.2040259f * d.y + 1.0572252f * d.z};
}
DEV float heroBalance(F4 t) {
    float s = sum(t);
    return s > 0.0f ? t.x / s : 0.0f;
}
DEV float dispersiveIor(float ior, float abbe, float l) {
    if (ior <= 1.0f + 1e-4f || abbe <= 0.0f) return fmaxf(ior, 1.0f);
    float lum = fmaxf(l * 0.001f, 1e-4f);
    float nFC = (ior - 1.0f) / fmaxf(abbe, 1e-6f);
    float invF2 = 1.0f / (0.4861327f * 0.4861327f);
    float invC2 = 1.0f / (0.6562725f * 0.6562725f);
    float B = nFC / (invF2 - invC2);
    float A = ior - B / (0.5875618f * 0.5875618f);
    return fmaxf(A + B / (lum * lum), 1.0f);
}
// src/shaders/sampling/wavelength.slang:21-47
DEV float wavelengthUnit(Rng& r, uint32_t sampleIndex) {
    float radInv = (float)__builtin_bitreverse32(sampleIndex) * (1.0f / 4294967296.0f);
    return fracf(radInv + rnd(r));
}

}  // namespace mrt

#include "dev_bsdf.h"
