/*
 * ORACLE -- test infrastructure only (see oracle_math.h). CPU restatement of
 * the reference's per-pixel path tracer, compiled twice: FP32 mirror
 * (functions suffixed _f32) and FP64 (suffixed _f64). Each function cites the
 * reference Slang / C it restates. Inputs are the packed scene buffers the
 * product exports (VKRT_SceneBuffers), i.e. the contents of the reference's
 * descriptor bindings (src/shaders/scene/resources.slang:7-59).
 *
 * Not restated (no counterpart needed for the configurations): material
 * textures, normal maps, alpha any-hit, sheen (its LTC table is not carried),
 * denoiser feature outputs, debug views other than NEE-only / BSDF-only.
 */
#include "oracle_math.h"

#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "../include/vkrt_mi355x.h"

#if defined(ORACLE_F64)
#define SFX(name) name##_f64
#else
#define SFX(name) name##_f32
#endif

#define PI_R R(3.14159265358979323846)
#define INV_PI_R R(0.31830988618379067154)

/* src/shaders/scene/constants.slang:6-15 */
#define RAY_T_MIN R(0.001)
#define RAY_T_MAX R(10000.0)
#define SHADOW_ORIGIN_OFFSET R(0.001)
#define SHADOW_DISTANCE_OFFSET R(0.002)
#define RR_MIN R(0.05)
#define RR_MAX R(0.95)
#define SAFE_NORMALIZE_EPS R(1e-12)
#define TANGENT_PARALLEL R(0.999)

/* ------------------------------------------------------------------ RNG -- */
/* src/shaders/sampling/random.slang:7-27 (policy 0) and a PCG32 stream per
 * path (policy 1). Both draw 24-bit uniforms as random.slang:16-19. */
typedef struct {
    uint32_t policy;
    uint32_t h;     /* hash state */
    uint64_t state; /* pcg32 */
    uint64_t inc;
} Rng;

static uint32_t hash32(uint32_t v) {
    v ^= v >> 16;
    v *= 0x7feb352du;
    v ^= v >> 15;
    v *= 0x846ca68bu;
    v ^= v >> 16;
    return v;
}

static uint64_t mix64(uint64_t z) {
    z += 0x9e3779b97f4a7c15ull;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

static uint32_t pcgNext(Rng* r) {
    uint64_t old = r->state;
    r->state = old * 6364136223846793005ull + r->inc;
    uint32_t xs = (uint32_t)(((old >> 18u) ^ old) >> 27u);
    uint32_t rot = (uint32_t)(old >> 59u);
    return (xs >> rot) | (xs << ((32u - rot) & 31u));
}

static Rng rngInit(uint32_t policy, uint64_t seed, int px, int py, uint32_t frame, uint32_t sampleIndex) {
    Rng r;
    memset(&r, 0, sizeof(r));
    r.policy = policy;
    if (policy == 0u) {
        uint32_t s = (uint32_t)px * 73856093u;
        s ^= (uint32_t)py * 19349663u;
        s ^= frame * 83492791u;
        s ^= sampleIndex * 2654435761u;
        r.h = hash32(s);
    } else {
        uint64_t stream = mix64((((uint64_t)(uint32_t)px) << 32) | (uint32_t)py) ^
                          mix64((((uint64_t)frame) << 32) | sampleIndex);
        r.inc = (stream << 1u) | 1u;
        r.state = 0u;
        pcgNext(&r);
        r.state += seed;
        pcgNext(&r);
    }
    return r;
}

static real randr(Rng* r) {
    uint32_t bits;
    if (r->policy == 0u) {
        r->h = hash32(r->h + 0x9e3779b9u);
        bits = r->h & 0x00ffffffu;
    } else {
        bits = pcgNext(r) >> 8u;
    }
    return (real)bits * (R(1.0) / R(16777216.0));
}

/* --------------------------------------------------------- scene access -- */
typedef struct {
    const VKRT_SceneBuffers* b;
    const SceneData* sd;
    /* BVH over world triangles (oracle-owned, binned SAH, independent of the GPU build) */
    const struct OBvh* bvh;
} Ctx;

/* -------------------------------------------------------------- camera -- */
typedef struct {
    v3 o, d;
    real tmin, tmax;
} Ray;

static v4 mulM4(const vkrt_m4 m, v4 v) {
    /* column-major m[col][row]; the shaders use -matrix-layout-column-major */
    return mk4(R(m[0][0]) * v.x + R(m[1][0]) * v.y + R(m[2][0]) * v.z + R(m[3][0]) * v.w,
               R(m[0][1]) * v.x + R(m[1][1]) * v.y + R(m[2][1]) * v.z + R(m[3][1]) * v.w,
               R(m[0][2]) * v.x + R(m[1][2]) * v.y + R(m[2][2]) * v.z + R(m[3][2]) * v.w,
               R(m[0][3]) * v.x + R(m[1][3]) * v.y + R(m[2][3]) * v.z + R(m[3][3]) * v.w);
}

/* src/shaders/camera/ray.slang:15-30 */
static Ray makePrimaryRay(const SceneData* sd, int px, int py, real jx, real jy) {
    real vx = (real)(px - (int)sd->viewportRect[0]) + R(0.5) + jx;
    real vy = (real)(py - (int)sd->viewportRect[1]) + R(0.5) + jy;
    real u = vx / (real)sd->viewportRect[2];
    real v = vy / (real)sd->viewportRect[3];
    real nx = u * R(2.0) - R(1.0);
    real ny = v * R(2.0) - R(1.0);
    v4 viewDir = mulM4(sd->projInverse, mk4(nx, ny, R(1.0), R(1.0)));
    v4 origin = mulM4(sd->viewInverse, mk4(R(0), R(0), R(0), R(1)));
    v4 dir = mulM4(sd->viewInverse, mk4(viewDir.x, viewDir.y, viewDir.z, R(0)));
    Ray r;
    r.o = mk3(origin.x, origin.y, origin.z);
    r.d = normalize3(mk3(dir.x, dir.y, dir.z));
    r.tmin = RAY_T_MIN;
    r.tmax = RAY_T_MAX;
    return r;
}

/* src/shaders/camera/ray.slang:32-39 */
static v3 safeNormalize(v3 v) {
    real l2 = dot3(v, v);
    if (l2 <= SAFE_NORMALIZE_EPS) return mk3(R(0), R(0), R(1));
    return scl3(v, RSQRT_(l2));
}

/* -------------------------------------------------------- intersection -- */
/* Watertight ray/triangle test (Woop, Benthin, Wald 2013), two-sided because
 * the reference traces with RAY_FLAG_NONE and no culling
 * (src/shaders/rt/queries/scene_query.slang:7-8). Reports t and the
 * barycentrics (u, v) of vertices 1 and 2, as BuiltInTriangleIntersectionAttributes. */
typedef struct {
    int kx, ky, kz;
    real sx, sy, sz;
} RayPre;

static RayPre rayPrecompute(v3 d) {
    RayPre p;
    real ad[3] = {FABS_(d.x), FABS_(d.y), FABS_(d.z)};
    real dd[3] = {d.x, d.y, d.z};
    int kz = 0;
    if (ad[1] > ad[kz]) kz = 1;
    if (ad[2] > ad[kz]) kz = 2;
    int kx = (kz + 1) % 3, ky = (kx + 1) % 3;
    if (dd[kz] < R(0)) {
        int t = kx;
        kx = ky;
        ky = t;
    }
    p.kx = kx;
    p.ky = ky;
    p.kz = kz;
    p.sx = dd[kx] / dd[kz];
    p.sy = dd[ky] / dd[kz];
    p.sz = R(1.0) / dd[kz];
    return p;
}

static int intersectTri(const float* tri, v3 o, const RayPre* p, real tmin, real tmax, real* tOut, real* uOut,
                        real* vOut) {
    real A[3] = {R(tri[0]) - o.x, R(tri[1]) - o.y, R(tri[2]) - o.z};
    real B[3] = {R(tri[3]) - o.x, R(tri[4]) - o.y, R(tri[5]) - o.z};
    real C[3] = {R(tri[6]) - o.x, R(tri[7]) - o.y, R(tri[8]) - o.z};
    real Ax = A[p->kx] - p->sx * A[p->kz];
    real Ay = A[p->ky] - p->sy * A[p->kz];
    real Bx = B[p->kx] - p->sx * B[p->kz];
    real By = B[p->ky] - p->sy * B[p->kz];
    real Cx = C[p->kx] - p->sx * C[p->kz];
    real Cy = C[p->ky] - p->sy * C[p->kz];
    real U = Cx * By - Cy * Bx;
    real V = Ax * Cy - Ay * Cx;
    real W = Bx * Ay - By * Ax;
    if (U == R(0) || V == R(0) || W == R(0)) {
        double CxBy = (double)Cx * (double)By, CyBx = (double)Cy * (double)Bx;
        double AxCy = (double)Ax * (double)Cy, AyCx = (double)Ay * (double)Cx;
        double BxAy = (double)Bx * (double)Ay, ByAx = (double)By * (double)Ax;
        U = (real)(CxBy - CyBx);
        V = (real)(AxCy - AyCx);
        W = (real)(BxAy - ByAx);
    }
    if ((U < R(0) || V < R(0) || W < R(0)) && (U > R(0) || V > R(0) || W > R(0))) return 0;
    real det = U + V + W;
    if (det == R(0)) return 0;
    real Az = p->sz * A[p->kz];
    real Bz = p->sz * B[p->kz];
    real Cz = p->sz * C[p->kz];
    real T = U * Az + V * Bz + W * Cz;
    real invDet = R(1.0) / det;
    real t = T * invDet;
    if (!(t > tmin) || !(t < tmax)) return 0;
    *tOut = t;
    *uOut = V * invDet;
    *vOut = W * invDet;
    return 1;
}

#include "oracle_bvh.h"

typedef struct {
    uint32_t mesh, prim, hit;
    real t, u, v;
    uint32_t global;
} Hit;

static int slab(const ONode* n, v3 o, v3 inv, real tmax) {
    real lo[3] = {R(n->lo[0]), R(n->lo[1]), R(n->lo[2])}, hi[3] = {R(n->hi[0]), R(n->hi[1]), R(n->hi[2])};
    real oo[3] = {o.x, o.y, o.z}, ii[3] = {inv.x, inv.y, inv.z};
    real t0 = R(0), t1 = tmax;
    for (int a = 0; a < 3; a++) {
        real ta = (lo[a] - oo[a]) * ii[a];
        real tb = (hi[a] - oo[a]) * ii[a];
        if (ta != ta) ta = -INFINITY; /* 0 * inf on a slab boundary */
        if (tb != tb) tb = INFINITY;
        real tn = rmin(ta, tb), tf = rmax(ta, tb);
        t0 = rmax(t0, tn);
        t1 = rmin(t1, tf * R(1.0000004));
    }
    return t0 <= t1;
}

/* Alpha test of a candidate hit (src/shaders/rt/alpha_test.slang:8-66) with the
 * build's draw: one uniform number per (ray key, triangle) from the lowbias32
 * mix, instead of the next number of the path's stream in the driver's
 * any-hit order. Ray key = mix(mix(path index) ^ salt); salt from the frame
 * number, the RNG seed, the depth and the ray kind (0 extension, 1 light
 * shadow, 2 environment shadow, 3 query). Vertex-colour alpha is interpolated
 * in float with weights (1-u-v, u, v). */
static uint32_t alphaSalt(uint32_t frame, uint64_t seed, uint32_t depth, uint32_t kind) {
    return hash32(frame * 0x9E3779B9u ^ (uint32_t)seed ^ (uint32_t)(seed >> 32) * 0x85EBCA6Bu ^
                  ((depth << 2) | kind) * 0xC2B2AE35u);
}
static uint32_t alphaRayKey(uint32_t salt, uint32_t pathIndex) { return hash32(hash32(pathIndex) ^ salt); }
static uint32_t rayKey(const Ctx* c, uint32_t pathIndex, uint32_t depth, uint32_t kind) {
    return alphaRayKey(alphaSalt(c->sd->frameNumber, c->b->rngSeed, depth, kind), pathIndex);
}

/* ------------------------------------------------ material textures -- */
/* Mip level 0, bilinear, per-axis wrap (the samplers of
 * src/core/scene/textures.c:141-178 sampled by
 * src/shaders/material/textures.slang:52-56), in float: texel (x, y) of a
 * texture is pool texel first + y * width + x, row 0 at coordinate v = 0. */
static int wrapTexel(int i, int n, uint32_t mode) {
    if (mode == VKRT_TEXTURE_WRAP_CLAMP_TO_EDGE) return i < 0 ? 0 : (i > n - 1 ? n - 1 : i);
    if (mode == VKRT_TEXTURE_WRAP_MIRRORED_REPEAT) {
        /* (size - 1) - mirror((i mod 2size) - size), mirror(a) = a >= 0 ? a : -(1 + a) */
        int a = (int)(((long long)i % (2LL * n) + 2LL * n) % (2LL * n)) - n;
        return (n - 1) - (a >= 0 ? a : -(1 + a));
    }
    return (int)(((long long)i % n + n) % n);
}
static float texelCoord(float uv, uint32_t size) {
    float x = uv * (float)size - 0.5f;
    x = fmaxf(x, -16777216.0f); /* a NaN coordinate becomes the lower bound */
    return fminf(x, 16777216.0f);
}
static void sampleTexture(const Ctx* c, uint32_t tex, float u, float v, uint32_t packedWrap, float out[4]) {
    const uint32_t* info = c->b->textureInfo + 4u * tex;
    const int w = (int)info[1], h = (int)info[2];
    const float x = texelCoord(u, info[1]), y = texelCoord(v, info[2]);
    const float fx = floorf(x), fy = floorf(y);
    const float ax = x - fx, ay = y - fy;
    const int xs[2] = {wrapTexel((int)fx, w, packedWrap & 0xffffu), wrapTexel((int)fx + 1, w, packedWrap & 0xffffu)};
    const int ys[2] = {wrapTexel((int)fy, h, packedWrap >> 16), wrapTexel((int)fy + 1, h, packedWrap >> 16)};
    for (int k = 0; k < 4; k++) {
        float t[2][2];
        for (int j = 0; j < 2; j++)
            for (int i = 0; i < 2; i++)
                t[j][i] = c->b->textureTexels[((size_t)info[0] + (size_t)ys[j] * (size_t)w + (size_t)xs[i]) * 4u + k];
        const float top = t[0][0] * (1.0f - ax) + t[0][1] * ax;
        const float bottom = t[1][0] * (1.0f - ax) + t[1][1] * ax;
        out[k] = top * (1.0f - ay) + bottom * ay;
    }
}
/* src/shaders/geometry/surface/interpolation.slang:46-60 (texture coordinates) */
typedef struct {
    float uv[2][2]; /* [set][u, v] */
} TexCoords;
static TexCoords hitTexCoords(const Ctx* c, uint32_t mesh, uint32_t prim, float u, float v) {
    const MeshInfo* m = &c->b->meshInfos[mesh];
    const uint32_t tb = m->indexBase + prim * 3u;
    const ShaderVertex* a = &c->b->vertices[c->b->indices[tb + 0] + m->vertexBase];
    const ShaderVertex* b = &c->b->vertices[c->b->indices[tb + 1] + m->vertexBase];
    const ShaderVertex* d = &c->b->vertices[c->b->indices[tb + 2] + m->vertexBase];
    const float w = 1.0f - u - v;
    TexCoords r;
    for (int k = 0; k < 2; k++) {
        r.uv[0][k] = a->texcoord0[k] * w + b->texcoord0[k] * u + d->texcoord0[k] * v;
        r.uv[1][k] = a->texcoord1[k] * w + b->texcoord1[k] * u + d->texcoord1[k] * v;
    }
    return r;
}
/* sampleMaterialTexture (textures.slang:58-74): coordinate set, scale,
 * rotation about the origin, offset; an empty slot returns the fallback. */
static void sampleMaterialSlot(const Ctx* c, const Material* m, uint32_t slot, const TexCoords* tc,
                               const float fallback[4], float out[4]) {
    const uint32_t idx[4] = {m->baseColorTextureIndex, m->metallicRoughnessTextureIndex, m->normalTextureIndex,
                             m->emissiveTextureIndex};
    const uint32_t wrap[4] = {m->baseColorTextureWrap, m->metallicRoughnessTextureWrap, m->normalTextureWrap,
                              m->emissiveTextureWrap};
    const float* xf[4] = {m->baseColorTextureTransform, m->metallicRoughnessTextureTransform,
                          m->normalTextureTransform, m->emissiveTextureTransform};
    if (idx[slot] == VKRT_INVALID_INDEX) {
        for (int k = 0; k < 4; k++) out[k] = fallback[k];
        return;
    }
    const int set = ((m->textureTexcoordSets >> (slot * 8u)) & 0xffu) == 1u;
    const float su = tc->uv[set][0] * xf[slot][0], sv = tc->uv[set][1] * xf[slot][1];
    const float sn = sinf(m->textureRotations[slot]), cs = cosf(m->textureRotations[slot]);
    sampleTexture(c, idx[slot], cs * su - sn * sv + xf[slot][2], sn * su + cs * sv + xf[slot][3], wrap[slot], out);
}
static int materialHasTextures(const Material* m) {
    return m->baseColorTextureIndex != VKRT_INVALID_INDEX || m->metallicRoughnessTextureIndex != VKRT_INVALID_INDEX ||
           m->normalTextureIndex != VKRT_INVALID_INDEX || m->emissiveTextureIndex != VKRT_INVALID_INDEX;
}

static int alphaAccept(const Ctx* c, uint32_t tri, real u, real v, uint32_t key) {
    const uint32_t word = c->b->triangleAlpha[tri];
    if (!(word >> 31)) return 1;
    const MeshInfo* mesh = &c->b->meshInfos[c->b->triangleMesh[tri]];
    const Material* m = &c->b->materials[mesh->materialIndex];
    const float k = 1.0f / 255.0f;
    const float fu = (float)u, fv = (float)v;
    const float w = 1.0f - fu - fv;
    const float ca = (float)(word & 0xffu) * k * w + (float)((word >> 8) & 0xffu) * k * fu +
                     (float)((word >> 16) & 0xffu) * k * fv;
    /* evaluateMaterialTextureAlpha (alpha_test.slang:20-26): MASK and BLEND
     * materials multiply in the base-colour texture's alpha */
    float ta = ca;
    if (word & 0x40000000u) {
        const uint32_t mi = c->b->triangleMesh[tri];
        uint32_t base = 0;
        for (uint32_t k = 0; k < mi; k++) base += c->b->meshInfos[k].indexCount / 3u;
        const TexCoords tc = hitTexCoords(c, mi, tri - base, fu, fv);
        const float one[4] = {1.0f, 1.0f, 1.0f, 1.0f};
        float bc[4];
        sampleMaterialSlot(c, m, VKRT_MATERIAL_TEXTURE_SLOT_BASE_COLOR, &tc, one, bc);
        ta = bc[3] * ca;
    }
    const float x = (float)(hash32(key ^ tri * 0x27D4EB2Fu) >> 8) * (1.0f / 16777216.0f);
    if (m->alphaMode == VKRT_MATERIAL_ALPHA_MODE_MASK) {
        if (ta < m->alphaCutoff) return 0;
        float mo = mesh->opacity * m->opacity * ca;
        mo = mo < 0.0f ? 0.0f : (mo > 1.0f ? 1.0f : mo);
        if (mo >= 1.0f) return 1;
        if (mo <= 0.0f) return 0;
        return x <= mo;
    }
    float op = mesh->opacity * m->opacity * ta;
    op = op < 0.0f ? 0.0f : (op > 1.0f ? 1.0f : op);
    const int blend = m->alphaMode == VKRT_MATERIAL_ALPHA_MODE_BLEND || m->opacity < 0.999f || mesh->opacity < 0.999f;
    if (!blend) return 1;
    if (op <= 0.0f) return 0;
    if (op >= 1.0f) return 1;
    return x <= op;
}
#define ALPHA_OK(c, i, u, v, key) (!(c)->b->triangleAlpha || alphaAccept((c), (i), (u), (v), (key)))

/* The bound a triangle's t must stay below to become the closest hit: the
 * closest hit is the smallest t and, among equal t, the lowest triangle index
 * (what the brute-force loop's strict comparison in index order gives), so a
 * tree's leaf order cannot change it. */
static real tieBound(real best, uint32_t i, uint32_t bestTri) {
#ifdef ORACLE_F64
    return (bestTri != UINT32_MAX && i < bestTri) ? nextafter(best, INFINITY) : best;
#else
    return (bestTri != UINT32_MAX && i < bestTri) ? nextafterf(best, INFINITY) : best;
#endif
}

static void traceClosest(const Ctx* c, Ray ray, Hit* h, uint32_t key) {
    h->hit = 0;
    h->mesh = VKRT_INVALID_INDEX;
    h->prim = VKRT_INVALID_INDEX;
    h->t = R(0);
    h->u = h->v = R(0);
    RayPre p = rayPrecompute(ray.d);
    real best = ray.tmax;
    uint32_t bestTri = UINT32_MAX;
    real bu = 0, bv = 0;
    const float* tris = c->b->worldTriangles;
    if (!c->bvh) {
        for (uint32_t i = 0; i < c->b->triangleCount; i++) {
            real t, u, v;
            if (intersectTri(&tris[(size_t)i * 9u], ray.o, &p, ray.tmin, best, &t, &u, &v) &&
                ALPHA_OK(c, i, u, v, key)) {
                best = t;
                bestTri = i;
                bu = u;
                bv = v;
            }
        }
    } else {
        v3 inv = mk3(R(1) / ray.d.x, R(1) / ray.d.y, R(1) / ray.d.z);
        uint32_t stack[128];
        int sp = 0;
        stack[sp++] = 0;
        while (sp) {
            const ONode* n = &c->bvh->nodes[stack[--sp]];
            if (!slab(n, ray.o, inv, best)) continue;
            if (n->left < 0) {
                for (uint32_t k = 0; k < n->count; k++) {
                    uint32_t i = c->bvh->order[n->first + k];
                    real t, u, v;
                    if (intersectTri(&tris[(size_t)i * 9u], ray.o, &p, ray.tmin, tieBound(best, i, bestTri), &t, &u,
                                     &v) &&
                        ALPHA_OK(c, i, u, v, key)) {
                        best = t;
                        bestTri = i;
                        bu = u;
                        bv = v;
                    }
                }
            } else {
                stack[sp++] = (uint32_t)n->left + 1u;
                stack[sp++] = (uint32_t)n->left;
            }
        }
    }
    if (bestTri != UINT32_MAX) {
        uint32_t mesh = c->b->triangleMesh[bestTri];
        /* mesh-local primitive: meshes are packed in order, so count back to the first triangle of the mesh */
        uint32_t first = c->b->meshInfos[mesh].indexBase / 3u;
        h->hit = 1;
        h->mesh = mesh;
        h->prim = bestTri - first;
        h->global = bestTri;
        h->t = best;
        h->u = bu;
        h->v = bv;
    }
}

/* Any hit within (tmin, tmax): the reference's shadow TraceRay with
 * ACCEPT_FIRST_HIT_AND_END_SEARCH (src/shaders/rt/queries/shadow_query.slang:8-28).
 * Returns the global triangle index of the accepted occluder or UINT32_MAX. */
static uint32_t traceAny(const Ctx* c, Ray ray, uint32_t key) {
    RayPre p = rayPrecompute(ray.d);
    const float* tris = c->b->worldTriangles;
    if (!c->bvh) {
        for (uint32_t i = 0; i < c->b->triangleCount; i++) {
            real t, u, v;
            if (intersectTri(&tris[(size_t)i * 9u], ray.o, &p, ray.tmin, ray.tmax, &t, &u, &v) &&
                ALPHA_OK(c, i, u, v, key))
                return i;
        }
        return UINT32_MAX;
    }
    v3 inv = mk3(R(1) / ray.d.x, R(1) / ray.d.y, R(1) / ray.d.z);
    uint32_t stack[128];
    int sp = 0;
    stack[sp++] = 0;
    while (sp) {
        const ONode* n = &c->bvh->nodes[stack[--sp]];
        if (!slab(n, ray.o, inv, ray.tmax)) continue;
        if (n->left < 0) {
            for (uint32_t k = 0; k < n->count; k++) {
                uint32_t i = c->bvh->order[n->first + k];
                real t, u, v;
                if (intersectTri(&tris[(size_t)i * 9u], ray.o, &p, ray.tmin, ray.tmax, &t, &u, &v) &&
                    ALPHA_OK(c, i, u, v, key))
                    return i;
            }
        } else {
            stack[sp++] = (uint32_t)n->left + 1u;
            stack[sp++] = (uint32_t)n->left;
        }
    }
    return UINT32_MAX;
}

#include "oracle_shading.inc"
