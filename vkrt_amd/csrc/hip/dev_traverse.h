// Ray traversal for gfx950: replaces the hardware TraceRay the reference
// relies on (src/shaders/rt/queries/scene_query.slang:10-42,
// shadow_query.slang:8-28). Closest-hit and any-hit queries over a 4-wide BVH
// with per-axis child boxes; triangles are two-sided (RAY_FLAG_NONE, no
// culling) and tested with the watertight algorithm of Woop, Benthin and Wald
// (JCGT 2013), reporting (t, u, v) with u, v the weights of vertices 1 and 2 as
// BuiltInTriangleIntersectionAttributes.
//
// Stack: kLdsStack entries per lane in LDS (bank-conflict free layout
// [entry][lane]) and kGlobalStack more in a per-thread global region, used only
// when a traversal gets deep; pushes beyond both are dropped and counted.
#pragma once

#include "dev_common.h"

namespace mrt {

constexpr int kLdsStack = 16;
constexpr int kGlobalStack = 112;
constexpr int kTraceBlock = 256;
// Hard cap on loop iterations of one traversal: a well-formed BVH over 2^27
// triangles needs far fewer; the cap only guarantees termination.
constexpr uint32_t kMaxTraversalSteps = 1u << 22;

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
    F3 oInv;     // o/d, so entry distances are lo*inv - oInv
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

// Slab test of the four children; near/far planes picked by ray octant.
DEV void boxes4(const BvhNode* node, const TraversalRay& r, float tmin, float tmax, float tn[4], bool hit[4]) {
    const float4* n4 = reinterpret_cast<const float4*>(node);
    float4 lx = n4[0], hx = n4[1], ly = n4[2], hy = n4[3], lz = n4[4], hz = n4[5];
    float4 nx = r.negX ? hx : lx, fx = r.negX ? lx : hx;
    float4 ny = r.negY ? hy : ly, fy = r.negY ? ly : hy;
    float4 nz = r.negZ ? hz : lz, fz = r.negZ ? lz : hz;
    const float nxa[4] = {nx.x, nx.y, nx.z, nx.w}, fxa[4] = {fx.x, fx.y, fx.z, fx.w};
    const float nya[4] = {ny.x, ny.y, ny.z, ny.w}, fya[4] = {fy.x, fy.y, fy.z, fy.w};
    const float nza[4] = {nz.x, nz.y, nz.z, nz.w}, fza[4] = {fz.x, fz.y, fz.z, fz.w};
#pragma unroll
    for (int i = 0; i < 4; i++) {
        float t0x = nxa[i] * r.inv.x - r.oInv.x, t1x = fxa[i] * r.inv.x - r.oInv.x;
        float t0y = nya[i] * r.inv.y - r.oInv.y, t1y = fya[i] * r.inv.y - r.oInv.y;
        float t0z = nza[i] * r.inv.z - r.oInv.z, t1z = fza[i] * r.inv.z - r.oInv.z;
        float tnear = fmaxf(fmaxf(t0x, t0y), fmaxf(t0z, tmin));
        float tfar = fminf(fminf(t1x, t1y), fminf(t1z, tmax)) * 1.0000004f;
        tn[i] = tnear;
        hit[i] = tnear <= tfar;
    }
}

struct TraversalStack {
    uint32_t* lds;   // entry i at lds[i * kTraceBlock]
    uint32_t* glob;  // entry i (>= kLdsStack) at glob[i - kLdsStack]
    int sp;
    uint32_t dropped;
    DEV void push(uint32_t ref) {
        if (sp < kLdsStack) {
            lds[sp * kTraceBlock] = ref;
        } else if (sp < kLdsStack + kGlobalStack) {
            glob[sp - kLdsStack] = ref;
        } else {
            dropped++;
            return;
        }
        sp++;
    }
    DEV uint32_t pop() {
        sp--;
        return sp < kLdsStack ? lds[sp * kTraceBlock] : glob[sp - kLdsStack];
    }
};

// Closest hit in (tmin, tmax). Children are visited nearest first.
DEV Hit traceClosest(const SceneView& s, const TraversalRay& r, float tmin, float tmax, TraversalStack& st,
                     uint32_t* nodeVisits = nullptr, uint32_t* triTests = nullptr) {
    Hit h;
    h.t = tmax;
    h.tri = kNoHit;
    h.u = h.v = 0.0f;
    if (s.nodeCount == 0) return h;
    st.sp = 0;
    uint32_t ref = 0;
    for (uint32_t step = 0;; step++) {
        if (step >= kMaxTraversalSteps) {
            st.dropped++;
            break;
        }
        if (!(ref & kLeafBit)) {
            if (nodeVisits) (*nodeVisits)++;
            const BvhNode* node = s.nodes + ref;
            float tn[4];
            bool hit[4];
            boxes4(node, r, tmin, h.t, tn, hit);
            uint4 c = node->child;
            uint32_t refs[4] = {c.x, c.y, c.z, c.w};
            // collect hit children, then order by entry distance (small sorting network)
            float d[4];
            uint32_t k[4];
            int n = 0;
#pragma unroll
            for (int i = 0; i < 4; i++) {
                hit[i] = hit[i] && refs[i] != kEmptyRef;
                d[i] = hit[i] ? tn[i] : INFINITY;
                k[i] = hit[i] ? refs[i] : kEmptyRef;
                n += hit[i] ? 1 : 0;
            }
#define MRT_CSWAP(a, b)                         \
    if (d[b] < d[a]) {                          \
        float td = d[a];                        \
        d[a] = d[b];                            \
        d[b] = td;                              \
        uint32_t tk = k[a];                     \
        k[a] = k[b];                            \
        k[b] = tk;                              \
    }
            MRT_CSWAP(0, 1)
            MRT_CSWAP(2, 3)
            MRT_CSWAP(0, 2)
            MRT_CSWAP(1, 3)
            MRT_CSWAP(1, 2)
#undef MRT_CSWAP
            if (n == 0) {
                if (st.sp == 0) break;
                ref = st.pop();
                continue;
            }
            for (int i = n - 1; i >= 1; i--) st.push(k[i]);
            ref = k[0];
        } else {
            uint32_t first = leafFirst(ref), count = leafCount(ref);
            for (uint32_t i = 0; i < count; i++) {
                if (triTests) (*triTests)++;
                const TriRecord tr = s.tris[first + i];
                float t, u, v;
                if (intersectTriangle(tr, r.o, r.pre, tmin, h.t, t, u, v)) {
                    h.t = t;
                    h.tri = __float_as_uint(tr.v0.w);
                    h.u = u;
                    h.v = v;
                }
            }
            if (st.sp == 0) break;
            ref = st.pop();
        }
    }
    return h;
}

// Any hit in (tmin, tmax): the first accepted intersection ends the search, as
// RAY_FLAG_ACCEPT_FIRST_HIT_AND_END_SEARCH. Returns the global triangle id or kNoHit.
DEV uint32_t traceAny(const SceneView& s, const TraversalRay& r, float tmin, float tmax, TraversalStack& st) {
    if (s.nodeCount == 0) return kNoHit;
    st.sp = 0;
    uint32_t ref = 0;
    for (uint32_t step = 0;; step++) {
        if (step >= kMaxTraversalSteps) {
            st.dropped++;
            break;
        }
        if (!(ref & kLeafBit)) {
            const BvhNode* node = s.nodes + ref;
            float tn[4];
            bool hit[4];
            boxes4(node, r, tmin, tmax, tn, hit);
            uint4 c = node->child;
            uint32_t refs[4] = {c.x, c.y, c.z, c.w};
            uint32_t next = kEmptyRef;
#pragma unroll
            for (int i = 0; i < 4; i++) {
                if (!hit[i] || refs[i] == kEmptyRef) continue;
                if (next == kEmptyRef)
                    next = refs[i];
                else
                    st.push(refs[i]);
            }
            if (next == kEmptyRef) {
                if (st.sp == 0) break;
                ref = st.pop();
                continue;
            }
            ref = next;
        } else {
            uint32_t first = leafFirst(ref), count = leafCount(ref);
            for (uint32_t i = 0; i < count; i++) {
                const TriRecord tr = s.tris[first + i];
                float t, u, v;
                if (intersectTriangle(tr, r.o, r.pre, tmin, tmax, t, u, v)) return __float_as_uint(tr.v0.w);
            }
            if (st.sp == 0) break;
            ref = st.pop();
        }
    }
    return kNoHit;
}

}  // namespace mrt
