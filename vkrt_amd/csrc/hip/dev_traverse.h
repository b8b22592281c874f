// Ray traversal for gfx950: replaces the hardware TraceRay the reference
// relies on (src/shaders/rt/queries/scene_query.slang:10-42,
// shadow_query.slang:8-28). Closest-hit and any-hit queries over an 8-wide
// BVH with quantised child boxes; triangles are two-sided (RAY_FLAG_NONE, no
// culling) and tested with the watertight algorithm of Woop, Benthin and Wald
// (JCGT 2013), reporting (t, u, v) with u, v the weights of vertices 1 and 2 as
// BuiltInTriangleIntersectionAttributes.
#pragma once

#include "dev_common.h"
#include "dev_texture.h"

namespace mrt {

constexpr int kLdsStack = 12;
constexpr int kGlobalStack = 116;
constexpr int kTraceBlock = 256;
// Hard cap on loop iterations of one traversal: a well-formed BVH over 2^27
// triangles needs far fewer; the cap only guarantees termination.
constexpr uint32_t kMaxTraversalSteps = 1u << 22;

// ------------------------------------------------------------ alpha test --
// src/shaders/rt/alpha_test.slang:8-66. The reference's any-hit shaders draw
// rand(rng) from the path's stream for every candidate hit the driver reports,
// in the driver's traversal order. Here the draw is a hash of (ray key,
// triangle): each (ray, triangle) pair gets one fixed uniform number, so the
// closest accepted hit does not depend on traversal order and the oracle can
// restate it, while every hit is still accepted with the reference's
// probability. Ray key = mix(mix(path index) ^ salt), salt from the frame,
// the RNG seed, the path depth and the ray kind (0 extension, 1 light shadow,
// 2 environment shadow, 3 ray query).
__host__ __device__ inline uint32_t alphaMix(uint32_t v) {
    v ^= v >> 16;
    v *= 0x7feb352du;
    v ^= v >> 15;
    v *= 0x846ca68bu;
    v ^= v >> 16;
    return v;
}
__host__ __device__ inline uint32_t alphaSalt(uint32_t frame, uint64_t seed, uint32_t depth, uint32_t kind) {
    return alphaMix(frame * 0x9E3779B9u ^ (uint32_t)seed ^ (uint32_t)(seed >> 32) * 0x85EBCA6Bu ^
                    ((depth << 2) | kind) * 0xC2B2AE35u);
}
__host__ __device__ inline uint32_t alphaRayKey(uint32_t salt, uint32_t pathIndex) {
    return alphaMix(alphaMix(pathIndex) ^ salt);
}

// Whether a candidate hit at barycentrics (u, v) of global triangle `tri`
// survives the alpha test. The alpha word holds the three vertex-colour alpha
// bytes, and bit 30 when the material's base-colour texture alpha takes part
// (MASK and BLEND materials with that texture, alpha_test.slang:20-26).
template <bool TEXA = true>
DEV bool alphaAccept(const SceneView& s, uint32_t tri, float u, float v, uint32_t key) {
    const uint32_t word = s.triAlpha[tri];
    if (!(word >> 31)) return true;
    const uint32_t meshIndex = s.triMesh[tri];
    const MeshInfo& mesh = s.meshes[meshIndex];
    const Material& m = s.materials[mesh.materialIndex];
    const float k = 1.0f / 255.0f;
    const float w = 1.0f - u - v;
    const float ca = (float)(word & 0xffu) * k * w + (float)((word >> 8) & 0xffu) * k * u +
                     (float)((word >> 16) & 0xffu) * k * v;
    float ta = ca;
    if (TEXA && (word & 0x40000000u) && s.texels)
        ta = sampleMaterialSlot(s, m, VKRT_MATERIAL_TEXTURE_SLOT_BASE_COLOR, hitTexCoords(s, tri, meshIndex, u, v),
                                f4(1.0f))
                 .w *
             ca;
    const float x = (float)(alphaMix(key ^ tri * 0x27D4EB2Fu) >> 8) * (1.0f / 16777216.0f);
    if (m.alphaMode == VKRT_MATERIAL_ALPHA_MODE_MASK) {
        if (ta < m.alphaCutoff) return false;
        const float mo = fminf(fmaxf(mesh.opacity * m.opacity * ca, 0.0f), 1.0f);
        if (mo >= 1.0f) return true;
        if (mo <= 0.0f) return false;
        return x <= mo;
    }
    const float op = fminf(fmaxf(mesh.opacity * m.opacity * ta, 0.0f), 1.0f);
    const bool blend = m.alphaMode == VKRT_MATERIAL_ALPHA_MODE_BLEND || m.opacity < 0.999f || mesh.opacity < 0.999f;
    if (!blend) return true;
    if (op <= 0.0f) return false;
    if (op >= 1.0f) return true;
    return x <= op;
}

struct RayPre {
    int kx, ky, kz;
    float sx, sy, sz;
};

DEV float comp(F3 v, int k) { return k == 0 ? v.x : (k == 1 ? v.y : v.z); }

DEV RayPre rayPrecompute(F3 d) {
    RayPre p;
    float ax = fabsf(d.x), ay = fabsf(d.y), az = fabsf(d.z);
    int kz = 0;
    if (ay > ax) kz = 1;
    if (az > (kz == 0 ? ax : ay)) kz = 2;
    int kx = kz == 2 ? 0 : kz + 1;
    int ky = kx == 2 ? 0 : kx + 1;
    float dz = comp(d, kz);
    if (dz < 0.0f) {
        int t = kx;
        kx = ky;
        ky = t;
    }
    p.kx = kx;
    p.ky = ky;
    p.kz = kz;
    p.sx = comp(d, kx) / dz;
    p.sy = comp(d, ky) / dz;
    p.sz = 1.0f / dz;
    return p;
}

// Watertight ray/triangle test. Evaluated without FMA contraction so it gives
// the same answer as an IEEE single-precision evaluation of the same formula.
DEV bool intersectTriangle(const TriRecord& tr, F3 o, const RayPre& p, float tmin, float tmax, float& tOut,
                           float& uOut, float& vOut) {
#pragma clang fp contract(off)
    F3 A = F3{tr.v0.x - o.x, tr.v0.y - o.y, tr.v0.z - o.z};
    F3 B = F3{tr.v1.x - o.x, tr.v1.y - o.y, tr.v1.z - o.z};
    F3 C = F3{tr.v2.x - o.x, tr.v2.y - o.y, tr.v2.z - o.z};
    float Akz = comp(A, p.kz), Bkz = comp(B, p.kz), Ckz = comp(C, p.kz);
    float Ax = comp(A, p.kx) - p.sx * Akz;
    float Ay = comp(A, p.ky) - p.sy * Akz;
    float Bx = comp(B, p.kx) - p.sx * Bkz;
    float By = comp(B, p.ky) - p.sy * Bkz;
    float Cx = comp(C, p.kx) - p.sx * Ckz;
    float Cy = comp(C, p.ky) - p.sy * Ckz;
    float U = Cx * By - Cy * Bx;
    float V = Ax * Cy - Ay * Cx;
    float W = Bx * Ay - By * Ax;
    if (U == 0.0f || V == 0.0f || W == 0.0f) {
        double CxBy = (double)Cx * (double)By, CyBx = (double)Cy * (double)Bx;
        double AxCy = (double)Ax * (double)Cy, AyCx = (double)Ay * (double)Cx;
        double BxAy = (double)Bx * (double)Ay, ByAx = (double)By * (double)Ax;
        U = (float)(CxBy - CyBx);
        V = (float)(AxCy - AyCx);
        W = (float)(BxAy - ByAx);
    }
    if ((U < 0.0f || V < 0.0f || W < 0.0f) && (U > 0.0f || V > 0.0f || W > 0.0f)) return false;
    float det = U + V + W;
    if (det == 0.0f) return false;
    float Az = p.sz * Akz;
    float Bz = p.sz * Bkz;
    float Cz = p.sz * Ckz;
    float T = U * Az + V * Bz + W * Cz;
    float invDet = 1.0f / det;
    float t = T * invDet;
    if (!(t > tmin) || !(t < tmax)) return false;
    tOut = t;
    uOut = V * invDet;
    vOut = W * invDet;
    return true;
}

struct TraversalRay {
    F3 o, d;
    F3 inv;      // 1/d
    F3 oInv;     // o/d, so entry distances are plane*inv - oInv
    bool negX, negY, negZ;
    RayPre pre;
};

DEV TraversalRay makeTraversalRay(F3 o, F3 d) {
    TraversalRay r;
    r.o = o;
    r.d = d;
    r.inv = F3{1.0f / d.x, 1.0f / d.y, 1.0f / d.z};
    r.oInv = F3{o.x * r.inv.x, o.y * r.inv.y, o.z * r.inv.z};
    r.negX = d.x < 0.0f;
    r.negY = d.y < 0.0f;
    r.negZ = d.z < 0.0f;
    r.pre = rayPrecompute(d);
    return r;
}

DEV float planeByte(uint32_t w0, uint32_t w1, int i) {
    return (float)(((i < 4 ? w0 : w1) >> ((i & 3) * 8)) & 0xFFu);
}

// The eight children of a node as registers: entry distances (misses at +inf)
// and references. Kept as vector components, never as an indexable array, so
// the compiler keeps them in VGPRs.
struct Children {
    float4 dLo, dHi;
    uint4 rLo, rHi;
    uint32_t mask;  // children whose box the ray overlaps
};

DEV void setComp(float4& v, int i, float x) {
    if (i == 0) v.x = x;
    else if (i == 1) v.y = x;
    else if (i == 2) v.z = x;
    else v.w = x;
}

// Slab test of the eight children of a node: decodes each child box, picks
// near and far planes by ray octant, and reports which children overlap
// [tmin, tmax] with their entry distances. The far distance is widened by a
// few ulp so that rounding in the decode and slab arithmetic never rejects a
// box the ray touches.
DEV Children boxes8(const BvhNode* node, const TraversalRay& r, float tmin, float tmax) {
    const uint4* w = reinterpret_cast<const uint4*>(node);
    const uint4 w0 = w[0], w1 = w[1], w2 = w[2], w3 = w[3];
    Children c;
    c.rLo = w[4];
    c.rHi = w[5];
    const float ox = __uint_as_float(w0.x), oy = __uint_as_float(w0.y), oz = __uint_as_float(w0.z);
    const uint32_t packed = w0.w;
    const float sx = __uint_as_float((packed & 0xFFu) << 23);
    const float sy = __uint_as_float(((packed >> 8) & 0xFFu) << 23);
    const float sz = __uint_as_float(((packed >> 16) & 0xFFu) << 23);
    const uint32_t count = packed >> 24;
    // near / far plane words per axis (lower planes in w1, w2.xy; upper in w2.zw, w3)
    const uint32_t nx0 = r.negX ? w2.z : w1.x, nx1 = r.negX ? w2.w : w1.y;
    const uint32_t fx0 = r.negX ? w1.x : w2.z, fx1 = r.negX ? w1.y : w2.w;
    const uint32_t ny0 = r.negY ? w3.x : w1.z, ny1 = r.negY ? w3.y : w1.w;
    const uint32_t fy0 = r.negY ? w1.z : w3.x, fy1 = r.negY ? w1.w : w3.y;
    const uint32_t nz0 = r.negZ ? w3.z : w2.x, nz1 = r.negZ ? w3.w : w2.y;
    const uint32_t fz0 = r.negZ ? w2.x : w3.z, fz1 = r.negZ ? w2.y : w3.w;
    c.mask = 0;
    // The plane code q decodes to origin + q * step; its slab distance
    // (plane - o) / d is q * (step / d) + (origin / d - o / d), one fma per
    // plane: step / d is exact (step is a power of two) and the decoded planes
    // contain the children without rounding (see quantizePlanes).
    const float ax = sx * r.inv.x, ay = sy * r.inv.y, az = sz * r.inv.z;
    const float bx = fmaf(ox, r.inv.x, -r.oInv.x), by = fmaf(oy, r.inv.y, -r.oInv.y),
                bz = fmaf(oz, r.inv.z, -r.oInv.z);
#pragma unroll
    for (int i = 0; i < 8; i++) {
        float t0x = fmaf(planeByte(nx0, nx1, i), ax, bx), t1x = fmaf(planeByte(fx0, fx1, i), ax, bx);
        float t0y = fmaf(planeByte(ny0, ny1, i), ay, by), t1y = fmaf(planeByte(fy0, fy1, i), ay, by);
        float t0z = fmaf(planeByte(nz0, nz1, i), az, bz), t1z = fmaf(planeByte(fz0, fz1, i), az, bz);
        float tnear = fmaxf(fmaxf(t0x, t0y), fmaxf(t0z, tmin));
        float tfar = fminf(fminf(t1x, t1y), fminf(t1z, tmax)) * 1.0000005f;
        bool hit = tnear <= tfar && (uint32_t)i < count;
        c.mask |= hit ? (1u << i) : 0u;
        if (i < 4)
            setComp(c.dLo, i, hit ? tnear : INFINITY);
        else
            setComp(c.dHi, i - 4, hit ? tnear : INFINITY);
    }
    return c;
}

// Sorts the children by entry distance (misses last) with a 19-comparator
// network for 8 inputs (Knuth, TAOCP 5.3.4).
DEV void sortChildren(Children& c) {
    float d0 = c.dLo.x, d1 = c.dLo.y, d2 = c.dLo.z, d3 = c.dLo.w, d4 = c.dHi.x, d5 = c.dHi.y, d6 = c.dHi.z,
          d7 = c.dHi.w;
    uint32_t r0 = c.rLo.x, r1 = c.rLo.y, r2 = c.rLo.z, r3 = c.rLo.w, r4 = c.rHi.x, r5 = c.rHi.y, r6 = c.rHi.z,
             r7 = c.rHi.w;
#define MRT_CE(a, b)                           \
    {                                          \
        bool sw = d##b < d##a;                 \
        float td = sw ? d##b : d##a;           \
        d##b = sw ? d##a : d##b;               \
        d##a = td;                             \
        uint32_t tr = sw ? r##b : r##a;        \
        r##b = sw ? r##a : r##b;               \
        r##a = tr;                             \
    }
    MRT_CE(0, 2) MRT_CE(1, 3) MRT_CE(4, 6) MRT_CE(5, 7)
    MRT_CE(0, 4) MRT_CE(1, 5) MRT_CE(2, 6) MRT_CE(3, 7)
    MRT_CE(0, 1) MRT_CE(2, 3) MRT_CE(4, 5) MRT_CE(6, 7)
    MRT_CE(2, 4) MRT_CE(3, 5)
    MRT_CE(1, 4) MRT_CE(3, 6)
    MRT_CE(1, 2) MRT_CE(3, 4) MRT_CE(5, 6)
#undef MRT_CE
    c.dLo = make_float4(d0, d1, d2, d3);
    c.dHi = make_float4(d4, d5, d6, d7);
    c.rLo = make_uint4(r0, r1, r2, r3);
    c.rHi = make_uint4(r4, r5, r6, r7);
}

// Per-lane traversal stack of (reference, entry distance) pairs: kLdsStack
// entries in LDS, laid out [entry][lane] so a wavefront's pushes hit distinct
// banks, and kGlobalStack more in a per-thread global region for deep
// traversals; pushes beyond both are dropped and counted.
struct TraversalStack {
    uint32_t* lds;   // reference i at lds[i * kTraceBlock], distance at lds[(kLdsStack + i) * kTraceBlock]
    uint32_t* glob;  // reference i (>= kLdsStack) at glob[i - kLdsStack], distance kGlobalStack words later
    int sp;
    uint32_t dropped;
    DEV void push(uint32_t ref, float d) {
        if (sp < kLdsStack) {
            lds[sp * kTraceBlock] = ref;
            lds[(kLdsStack + sp) * kTraceBlock] = __float_as_uint(d);
        } else if (sp < kLdsStack + kGlobalStack) {
            glob[sp - kLdsStack] = ref;
            glob[kGlobalStack + sp - kLdsStack] = __float_as_uint(d);
        } else {
            dropped++;
            return;
        }
        sp++;
    }
    DEV void pop(uint32_t& ref, float& d) {
        sp--;
        if (sp < kLdsStack) {
            ref = lds[sp * kTraceBlock];
            d = __uint_as_float(lds[(kLdsStack + sp) * kTraceBlock]);
        } else {
            ref = glob[sp - kLdsStack];
            d = __uint_as_float(glob[kGlobalStack + sp - kLdsStack]);
        }
    }
};
constexpr int kLdsStackWords = 2 * kLdsStack * kTraceBlock;
constexpr int kGlobalStackWords = 2 * kGlobalStack;

// Closest-hit traversal state of one ray.
struct ClosestState {
    Hit h;
    uint32_t ref;    // next node or leaf to visit
    uint32_t steps;  // visits so far (bounded by kMaxTraversalSteps)
};

DEV void closestBegin(ClosestState& cs, TraversalStack& st, float tmax) {
    cs.h.t = tmax;
    cs.h.tri = kNoHit;
    cs.h.u = cs.h.v = 0.0f;
    cs.ref = 0;
    cs.steps = 0;
    st.sp = 0;
}

// One node or leaf visit of a closest-hit traversal in (tmin, h.t). The
// nearest overlapping child is visited next and the others are pushed farthest
// first; a popped entry whose entry distance lies beyond the closest hit found
// so far is skipped. Returns false once the traversal has finished.
// ALPHA = false compiles the alpha test out (scenes where every mesh is opaque);
// TEXA = false compiles out its texture-alpha lookup (no material textures).
template <bool ALPHA = true, bool TEXA = true>
DEV bool closestStep(const SceneView& s, const TraversalRay& r, float tmin, ClosestState& cs, TraversalStack& st,
                     uint32_t alphaKey, uint32_t* nodeVisits = nullptr, uint32_t* triTests = nullptr) {
    if (++cs.steps > kMaxTraversalSteps) {
        st.dropped++;
        return false;
    }
    bool popNext = true;
    if (!(cs.ref & kLeafBit)) {
        if (nodeVisits) (*nodeVisits)++;
        Children c = boxes8(s.nodes + cs.ref, r, tmin, cs.h.t);
        const int hits = __popc(c.mask);
        if (hits > 0) {
            sortChildren(c);
            // visit the nearest next; push the others farthest first
            cs.ref = c.rLo.x;
            if (hits > 7) st.push(c.rHi.w, c.dHi.w);
            if (hits > 6) st.push(c.rHi.z, c.dHi.z);
            if (hits > 5) st.push(c.rHi.y, c.dHi.y);
            if (hits > 4) st.push(c.rHi.x, c.dHi.x);
            if (hits > 3) st.push(c.rLo.w, c.dLo.w);
            if (hits > 2) st.push(c.rLo.z, c.dLo.z);
            if (hits > 1) st.push(c.rLo.y, c.dLo.y);
            popNext = false;
        }
    } else {
        uint32_t first = leafFirst(cs.ref), count = leafCount(cs.ref);
        for (uint32_t i = 0; i < count; i++) {
            if (triTests) (*triTests)++;
            const TriRecord tr = s.tris[first + i];
            const uint32_t tri = __float_as_uint(tr.v0.w);
            // the closest hit is the smallest t and, among equal t (two triangles
            // sharing an edge), the lowest triangle id, whatever order the leaves
            // are visited in: a lower id may also take the current t itself
            const float tcap = (cs.h.tri != kNoHit && tri < cs.h.tri)
                                   ? __uint_as_float(__float_as_uint(cs.h.t) + 1u) : cs.h.t;
            float t, u, v;
            if (intersectTriangle(tr, r.o, r.pre, tmin, tcap, t, u, v)) {
                if (ALPHA && s.triAlpha && !alphaAccept<TEXA>(s, tri, u, v, alphaKey)) continue;
                cs.h.t = t;
                cs.h.tri = tri;
                cs.h.u = u;
                cs.h.v = v;
            }
        }
    }
    if (popNext) {
        while (st.sp > 0) {
            float d;
            st.pop(cs.ref, d);
            if (d <= cs.h.t) return true;
        }
        return false;
    }
    return true;
}

// Whether a ray can reach the scene's (padded) bounding box within
// [tmin, tmax]. Conservative: slab distances are computed with the ray's
// reciprocal direction, a NaN slab (an axis-parallel ray on a box plane)
// constrains nothing, and the far distance is widened.
DEV bool reachesSceneBox(const SceneView& s, const TraversalRay& r, float tmin, float tmax) {
    const float ax = (s.boxLo[0] - r.o.x) * r.inv.x, bx = (s.boxHi[0] - r.o.x) * r.inv.x;
    const float ay = (s.boxLo[1] - r.o.y) * r.inv.y, by = (s.boxHi[1] - r.o.y) * r.inv.y;
    const float az = (s.boxLo[2] - r.o.z) * r.inv.z, bz = (s.boxHi[2] - r.o.z) * r.inv.z;
    const float tnear = fmaxf(fmaxf(fminf(ax, bx), fminf(ay, by)), fmaxf(fminf(az, bz), tmin));
    const float tfar = fminf(fminf(fmaxf(ax, bx), fmaxf(ay, by)), fminf(fmaxf(az, bz), tmax));
    return tnear <= tfar * 1.00001f || !(tfar == tfar);
}

// Closest hit in (tmin, tmax).
DEV Hit traceClosest(const SceneView& s, const TraversalRay& r, float tmin, float tmax, TraversalStack& st,
                     uint32_t alphaKey, uint32_t* nodeVisits = nullptr, uint32_t* triTests = nullptr) {
    ClosestState cs;
    closestBegin(cs, st, tmax);
    if (s.nodeCount == 0) return cs.h;
    while (closestStep<true>(s, r, tmin, cs, st, alphaKey, nodeVisits, triTests)) {
    }
    return cs.h;
}

// Any hit in (tmin, tmax): the first accepted intersection ends the search, as
// RAY_FLAG_ACCEPT_FIRST_HIT_AND_END_SEARCH. Returns the global triangle id or kNoHit.
DEV uint32_t traceAny(const SceneView& s, const TraversalRay& r, float tmin, float tmax, TraversalStack& st,
                      uint32_t alphaKey) {
    if (s.nodeCount == 0) return kNoHit;
    st.sp = 0;
    uint32_t ref = 0;
    for (uint32_t step = 0;; step++) {
        if (step >= kMaxTraversalSteps) {
            st.dropped++;
            break;
        }
        bool popNext = true;
        if (!(ref & kLeafBit)) {
            const Children c = boxes8(s.nodes + ref, r, tmin, tmax);
            if (c.mask != 0u) {
                // visit the lowest overlapping slot next, push the others
                const uint32_t m = c.mask;
                const int first = __ffs(m) - 1;
                ref = first == 0 ? c.rLo.x : first == 1 ? c.rLo.y : first == 2 ? c.rLo.z : first == 3 ? c.rLo.w
                    : first == 4 ? c.rHi.x : first == 5 ? c.rHi.y : first == 6 ? c.rHi.z : c.rHi.w;
                const uint32_t rest = m & (m - 1u);
                if (rest & 0x02u) st.push(c.rLo.y, 0.0f);
                if (rest & 0x04u) st.push(c.rLo.z, 0.0f);
                if (rest & 0x08u) st.push(c.rLo.w, 0.0f);
                if (rest & 0x10u) st.push(c.rHi.x, 0.0f);
                if (rest & 0x20u) st.push(c.rHi.y, 0.0f);
                if (rest & 0x40u) st.push(c.rHi.z, 0.0f);
                if (rest & 0x80u) st.push(c.rHi.w, 0.0f);
                popNext = false;
            }
        } else {
            uint32_t first = leafFirst(ref), count = leafCount(ref);
            for (uint32_t i = 0; i < count; i++) {
                const TriRecord tr = s.tris[first + i];
                float t, u, v;
                if (intersectTriangle(tr, r.o, r.pre, tmin, tmax, t, u, v)) {
                    const uint32_t tri = __float_as_uint(tr.v0.w);
                    if (!s.triAlpha || alphaAccept(s, tri, u, v, alphaKey)) return tri;
                }
            }
        }
        if (popNext) {
            if (st.sp == 0) break;
            float d;
            st.pop(ref, d);
        }
    }
    return kNoHit;
}

}  // namespace mrt
