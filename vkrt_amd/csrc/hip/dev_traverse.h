// Ray traversal for gfx950: replaces the hardware TraceRay the reference
// relies on (src/shaders/rt/queries/scene_query.slang:10-42,
// shadow_query.slang:8-28). Closest-hit and any-hit queries over an 8-wide
// BVH with quantised child boxes; triangles are two-sided (RAY_FLAG_NONE, no
// culling) and tested with the watertight algorithm of Woop, Benthin and Wald
// (JCGT 2013), reporting (t, u, v) with u, v the weights of vertices 1 and 2 as
// BuiltInTriangleIntersectionAttributes.
#pragma once

#include "dev_common.h"

namespace mrt {

constexpr int kLdsStack = 12;
constexpr int kGlobalStack = 116;
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

// Slab test of the eight children of a node: decodes each child box, picks
// near and far planes by ray octant, and returns the mask of children whose
// box overlaps [tmin, tmax] with their entry distances in tn and their
// references in ch. The far distance is widened by 4 ulp so that rounding in
// the decode and slab arithmetic never rejects a box the ray touches.
DEV uint32_t boxes8(const BvhNode* node, const TraversalRay& r, float tmin, float tmax, float tn[8], uint32_t ch[8]) {
    const uint4* w = reinterpret_cast<const uint4*>(node);
    const uint4 w0 = w[0], w1 = w[1], w2 = w[2], w3 = w[3], w4 = w[4], w5 = w[5];
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
    ch[0] = w4.x;
    ch[1] = w4.y;
    ch[2] = w4.z;
    ch[3] = w4.w;
    ch[4] = w5.x;
    ch[5] = w5.y;
    ch[6] = w5.z;
    ch[7] = w5.w;
    uint32_t mask = 0;
#pragma unroll
    for (int i = 0; i < 8; i++) {
        float bnx = fmaf(planeByte(nx0, nx1, i), sx, ox), bfx = fmaf(planeByte(fx0, fx1, i), sx, ox);
        float bny = fmaf(planeByte(ny0, ny1, i), sy, oy), bfy = fmaf(planeByte(fy0, fy1, i), sy, oy);
        float bnz = fmaf(planeByte(nz0, nz1, i), sz, oz), bfz = fmaf(planeByte(fz0, fz1, i), sz, oz);
        float t0x = fmaf(bnx, r.inv.x, -r.oInv.x), t1x = fmaf(bfx, r.inv.x, -r.oInv.x);
        float t0y = fmaf(bny, r.inv.y, -r.oInv.y), t1y = fmaf(bfy, r.inv.y, -r.oInv.y);
        float t0z = fmaf(bnz, r.inv.z, -r.oInv.z), t1z = fmaf(bfz, r.inv.z, -r.oInv.z);
        float tnear = fmaxf(fmaxf(t0x, t0y), fmaxf(t0z, tmin));
        float tfar = fminf(fminf(t1x, t1y), fminf(t1z, tmax)) * 1.0000005f;
        tn[i] = tnear;
        mask |= (tnear <= tfar && (uint32_t)i < count) ? (1u << i) : 0u;
    }
    return mask;
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

// Child i of the mask with the smallest (largest) entry distance, chosen with
// unrolled selects so tn and ch stay in registers.
DEV int nearestChild(uint32_t mask, const float tn[8]) {
    float best = INFINITY;
    int bi = 0;
#pragma unroll
    for (int i = 0; i < 8; i++) {
        bool take = ((mask >> i) & 1u) && tn[i] < best;
        best = take ? tn[i] : best;
        bi = take ? i : bi;
    }
    return bi;
}
DEV int farthestChild(uint32_t mask, const float tn[8]) {
    float best = -INFINITY;
    int bi = 0;
#pragma unroll
    for (int i = 0; i < 8; i++) {
        bool take = ((mask >> i) & 1u) && tn[i] >= best;
        best = take ? tn[i] : best;
        bi = take ? i : bi;
    }
    return bi;
}
DEV uint32_t selectRef(const uint32_t ch[8], int k) {
    uint32_t r = ch[0];
#pragma unroll
    for (int i = 1; i < 8; i++) r = k == i ? ch[i] : r;
    return r;
}
DEV float selectDist(const float tn[8], int k) {
    float d = tn[0];
#pragma unroll
    for (int i = 1; i < 8; i++) d = k == i ? tn[i] : d;
    return d;
}

// Closest hit in (tmin, tmax). The nearest overlapping child is visited next
// and the others are pushed farthest first; a popped entry whose entry
// distance lies beyond the closest hit found so far is skipped.
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
        bool popNext = true;
        if (!(ref & kLeafBit)) {
            if (nodeVisits) (*nodeVisits)++;
            float tn[8];
            uint32_t ch[8];
            uint32_t mask = boxes8(s.nodes + ref, r, tmin, h.t, tn, ch);
            if (mask != 0u) {
                int ni = nearestChild(mask, tn);
                ref = selectRef(ch, ni);
                mask &= ~(1u << ni);
                while (mask != 0u) {
                    int fi = farthestChild(mask, tn);
                    st.push(selectRef(ch, fi), selectDist(tn, fi));
                    mask &= ~(1u << fi);
                }
                popNext = false;
            }
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
        }
        if (popNext) {
            bool found = false;
            while (st.sp > 0) {
                float d;
                st.pop(ref, d);
                if (d <= h.t) {
                    found = true;
                    break;
                }
            }
            if (!found) break;
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
        bool popNext = true;
        if (!(ref & kLeafBit)) {
            float tn[8];
            uint32_t ch[8];
            uint32_t mask = boxes8(s.nodes + ref, r, tmin, tmax, tn, ch);
            if (mask != 0u) {
                int ni = nearestChild(mask, tn);
                ref = selectRef(ch, ni);
                mask &= ~(1u << ni);
                while (mask != 0u) {
                    int k = __ffs(mask) - 1;
                    st.push(selectRef(ch, k), 0.0f);
                    mask &= mask - 1u;
                }
                popNext = false;
            }
        } else {
            uint32_t first = leafFirst(ref), count = leafCount(ref);
            for (uint32_t i = 0; i < count; i++) {
                const TriRecord tr = s.tris[first + i];
                float t, u, v;
                if (intersectTriangle(tr, r.o, r.pre, tmin, tmax, t, u, v)) return __float_as_uint(tr.v0.w);
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
