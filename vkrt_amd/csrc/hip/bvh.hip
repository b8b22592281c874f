// On-GPU BVH build for gfx950, replacing the driver-built BLAS/TLAS of the
// reference (src/core/render/accel/blas.c:222-262, tlas.c:535-559).
//
//  1. per-triangle bounds and the centroid bounds of the scene
//  2. 63-bit Morton codes (21 bits per axis) of triangle centroids
//  3. radix sort of (code, triangle) pairs (rocPRIM)
//  4. binary radix tree over the sorted codes (Karras, HPG 2012), each node
//     covering a contiguous range of sorted triangles
//  5. bottom-up bounds with one atomic arrival flag per node
//  6. top-down collapse into 8-wide nodes: a node is opened greedily by
//     splitting its largest-area child until eight children remain; subtrees
//     of at most kLeafMaxTris triangles become leaves (level-synchronous);
//     child boxes are quantised to 8 bits per plane, conservatively
//  7. triangles reordered into leaf order as 48-byte records
#include <cstring>
#include <rocprim/device/device_radix_sort.hpp>

#include <algorithm>
#include <chrono>
#include <vector>

#include "mrt_internal.hpp"

namespace mrt {

constexpr int kB = 256;

struct Aabb {
    float lo[3], hi[3];
};

__device__ __forceinline__ uint32_t floatToOrdered(float f) {
    uint32_t u = __float_as_uint(f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float orderedToFloat(uint32_t u) {
    return __uint_as_float((u & 0x80000000u) ? (u & 0x7fffffffu) : ~u);
}

// per-triangle box, centroid bounds reduced into ordered-int min/max
__global__ __launch_bounds__(kB) void kTriBounds(const float* tris9, uint32_t n, Aabb* boxes, uint32_t* centroidBounds) {
    uint32_t i = blockIdx.x * kB + threadIdx.x;
    float c[3] = {INFINITY, INFINITY, INFINITY}, C[3] = {-INFINITY, -INFINITY, -INFINITY};
    if (i < n) {
        const float* t = tris9 + (size_t)i * 9u;
        Aabb b;
        for (int a = 0; a < 3; a++) {
            b.lo[a] = fminf(fminf(t[a], t[3 + a]), t[6 + a]);
            b.hi[a] = fmaxf(fmaxf(t[a], t[3 + a]), t[6 + a]);
            float m = 0.5f * (b.lo[a] + b.hi[a]);
            c[a] = m;
            C[a] = m;
        }
        boxes[i] = b;
    }
    __shared__ uint32_t red[6][kB];
    for (int a = 0; a < 3; a++) {
        red[a][threadIdx.x] = floatToOrdered(c[a]);
        red[3 + a][threadIdx.x] = floatToOrdered(C[a]);
    }
    __syncthreads();
    for (int s = kB / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) {
            for (int a = 0; a < 3; a++) {
                red[a][threadIdx.x] = min(red[a][threadIdx.x], red[a][threadIdx.x + s]);
                red[3 + a][threadIdx.x] = max(red[3 + a][threadIdx.x], red[3 + a][threadIdx.x + s]);
            }
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        for (int a = 0; a < 3; a++) {
            atomicMin(&centroidBounds[a], red[a][0]);
            atomicMax(&centroidBounds[3 + a], red[3 + a][0]);
        }
    }
}

__device__ __forceinline__ uint64_t spread21(uint64_t v) {
    v &= 0x1fffffull;
    v = (v | (v << 32)) & 0x1f00000000ffffull;
    v = (v | (v << 16)) & 0x1f0000ff0000ffull;
    v = (v | (v << 8)) & 0x100f00f00f00f00full;
    v = (v | (v << 4)) & 0x10c30c30c30c30c3ull;
    v = (v | (v << 2)) & 0x1249249249249249ull;
    return v;
}

__global__ __launch_bounds__(kB) void kMorton(const Aabb* boxes, uint32_t n, const uint32_t* centroidBounds,
                                             uint64_t* keys, uint32_t* values) {
    uint32_t i = blockIdx.x * kB + threadIdx.x;
    if (i >= n) return;
    uint64_t code = 0;
    for (int a = 0; a < 3; a++) {
        float lo = orderedToFloat(centroidBounds[a]);
        float hi = orderedToFloat(centroidBounds[3 + a]);
        float ext = hi - lo;
        float m = 0.5f * (boxes[i].lo[a] + boxes[i].hi[a]);
        float t = ext > 0.0f ? (m - lo) / ext : 0.5f;
        uint64_t q = (uint64_t)fminf(fmaxf(t * 2097152.0f, 0.0f), 2097151.0f);
        code |= spread21(q) << (2 - a);
    }
    keys[i] = code;
    values[i] = i;
}

// Binary radix tree (Karras 2012). Internal node i in [0, n-1); leaves are
// encoded as (n - 1 + sortedIndex). range[node] = first/last sorted index.
__device__ __forceinline__ int delta(const uint64_t* keys, int n, int i, int j) {
    if (j < 0 || j >= n) return -1;
    uint64_t a = keys[i], b = keys[j];
    if (a == b) return 64 + __clz((uint32_t)(i ^ j));
    return __clzll(a ^ b);
}

__global__ __launch_bounds__(kB) void kRadixTree(const uint64_t* keys, int n, int2* children, int* parent, int2* range) {
    int i = blockIdx.x * kB + threadIdx.x;
    if (i >= n - 1) return;
    int d = (delta(keys, n, i, i + 1) - delta(keys, n, i, i - 1)) >= 0 ? 1 : -1;
    int dMin = delta(keys, n, i, i - d);
    int lMax = 2;
    while (delta(keys, n, i, i + lMax * d) > dMin) lMax <<= 1;
    int l = 0;
    for (int t = lMax >> 1; t >= 1; t >>= 1)
        if (delta(keys, n, i, i + (l + t) * d) > dMin) l += t;
    int j = i + l * d;
    int dNode = delta(keys, n, i, j);
    int s = 0;
    int div = 2;
    int t = (l + div - 1) / div;
    while (true) {
        if (delta(keys, n, i, i + (s + t) * d) > dNode) s += t;
        if (t == 1) break;
        div <<= 1;
        t = (l + div - 1) / div;
    }
    int gamma = i + s * d + min(d, 0);
    int first = min(i, j), last = max(i, j);
    int left = (first == gamma) ? (n - 1 + gamma) : gamma;
    int right = (last == gamma + 1) ? (n - 1 + gamma + 1) : gamma + 1;
    children[i] = make_int2(left, right);
    range[i] = make_int2(first, last);
    parent[left] = i;
    parent[right] = i;
}

// bottom-up bounds: leaves start, the second arrival at a node merges
__global__ __launch_bounds__(kB) void kRefit(const Aabb* triBoxes, const uint32_t* sortedTri, int n, const int2* children,
                                            const int* parent, Aabb* nodeBoxes, uint32_t* arrival) {
    int i = blockIdx.x * kB + threadIdx.x;
    if (i >= n) return;
    int node = n - 1 + i;
    nodeBoxes[node] = triBoxes[sortedTri[i]];
    int p = parent[node];
    while (p >= 0) {
        __threadfence();
        if (atomicAdd(&arrival[p], 1u) == 0u) return;
        __threadfence();
        int2 c = children[p];
        // the sibling's box was written by another thread: read it through L2 (coherent loads)
        const volatile Aabb* va = &nodeBoxes[c.x];
        const volatile Aabb* vb = &nodeBoxes[c.y];
        Aabb m;
        for (int k = 0; k < 3; k++) {
            m.lo[k] = fminf(va->lo[k], vb->lo[k]);
            m.hi[k] = fmaxf(va->hi[k], vb->hi[k]);
        }
        nodeBoxes[p] = m;
        p = parent[p];
    }
}

__device__ __forceinline__ float halfArea(const Aabb& b) {
    float dx = fmaxf(b.hi[0] - b.lo[0], 0.0f), dy = fmaxf(b.hi[1] - b.lo[1], 0.0f), dz = fmaxf(b.hi[2] - b.lo[2], 0.0f);
    return dx * dy + dy * dz + dz * dx;
}

struct CollapseItem {
    int binNode;   // binary node (internal index or n-1+leaf)
    uint32_t wide;  // wide node slot to fill
};

__device__ __forceinline__ int rangeCount(int node, int n, const int2* range) {
    if (node >= n - 1) return 1;
    int2 r = range[node];
    return r.y - r.x + 1;
}
__device__ __forceinline__ int rangeFirst(int node, int n, const int2* range) {
    return node >= n - 1 ? node - (n - 1) : range[node].x;
}

// Quantisation of one axis of a node: the smallest power-of-two step with
// which 255 steps from the node's lower plane reach its upper plane, checked
// with the decode arithmetic of the traversal (fmaf(q, step, origin)).
__device__ uint32_t nodeExponent(float lo, float hi) {
    float ext = hi - lo;
    int e = 1;
    if (ext > 0.0f) {
        int ex;
        frexpf(ext / 255.0f, &ex);  // ext/255 = m * 2^ex, m in [0.5, 1): 2^ex >= ext/255
        e = min(max(ex + 127, 1), 254);
    }
    while (e < 254 && fmaf(255.0f, __uint_as_float((uint32_t)e << 23), lo) < hi) e++;
    return (uint32_t)e;
}
// Lower and upper plane codes of a child interval [clo, chi] inside the node:
// decode(qlo) <= clo and decode(qhi) >= chi exactly.
__device__ void quantizePlanes(float origin, float step, float clo, float chi, uint32_t& qlo, uint32_t& qhi) {
    float a = floorf((clo - origin) / step), b = ceilf((chi - origin) / step);
    int l = (int)fminf(fmaxf(a, 0.0f), 255.0f), h = (int)fminf(fmaxf(b, 0.0f), 255.0f);
    while (l > 0 && fmaf((float)l, step, origin) > clo) l--;
    while (h < 255 && fmaf((float)h, step, origin) < chi) h++;
    qlo = (uint32_t)l;
    qhi = (uint32_t)h;
}

__global__ __launch_bounds__(kB) void kCollapse(const CollapseItem* in, const uint32_t* inCount, CollapseItem* out,
                                               uint32_t* outCount, uint32_t* nodeAlloc, int n, const int2* children,
                                               const int2* range, const Aabb* nodeBoxes, BvhNode* nodes,
                                               uint32_t nodeCapacity, uint32_t* overflow) {
    uint32_t w = blockIdx.x * kB + threadIdx.x;
    if (w >= *inCount) return;
    CollapseItem it = in[w];
    int slot[kBvhWidth];
    int cnt = 0;
    if (it.binNode >= n - 1) {
        slot[cnt++] = it.binNode;
    } else {
        int2 c = children[it.binNode];
        slot[cnt++] = c.x;
        slot[cnt++] = c.y;
        while (cnt < kBvhWidth) {
            // open the largest child that is an internal node with too many triangles for a leaf
            int best = -1;
            float bestArea = -1.0f;
            for (int k = 0; k < cnt; k++) {
                int nd = slot[k];
                if (nd >= n - 1 || rangeCount(nd, n, range) <= (int)kLeafMaxTris) continue;
                float a = halfArea(nodeBoxes[nd]);
                if (a > bestArea) {
                    bestArea = a;
                    best = k;
                }
            }
            if (best < 0) break;
            int2 c2 = children[slot[best]];
            slot[best] = c2.x;
            slot[cnt++] = c2.y;
        }
    }
    const Aabb P = nodeBoxes[it.binNode];
    BvhNode node;
    node.ox = P.lo[0];
    node.oy = P.lo[1];
    node.oz = P.lo[2];
    uint32_t e[3];
    float step[3];
    for (int a = 0; a < 3; a++) {
        e[a] = nodeExponent(P.lo[a], P.hi[a]);
        step[a] = __uint_as_float(e[a] << 23);
    }
    for (int k = 0; k < 6; k++) node.qlo[k] = node.qhi[k] = 0u;
    uint32_t used = 0;
    for (int k = 0; k < cnt; k++) {
        int nd = slot[k];
        uint32_t ref;
        int count = rangeCount(nd, n, range);
        if (count <= (int)kLeafMaxTris) {
            ref = kLeafBit | ((uint32_t)(count - 1) << 27) | (uint32_t)rangeFirst(nd, n, range);
        } else {
            uint32_t child = atomicAdd(nodeAlloc, 1u);
            if (child >= nodeCapacity) {
                atomicAdd(overflow, 1u);
                continue;
            }
            ref = child;
            uint32_t o = atomicAdd(outCount, 1u);
            out[o] = CollapseItem{nd, child};
        }
        const Aabb b = nodeBoxes[nd];
        for (int a = 0; a < 3; a++) {
            uint32_t ql, qh;
            quantizePlanes(P.lo[a], step[a], b.lo[a], b.hi[a], ql, qh);
            node.qlo[a * 2 + (used >> 2)] |= ql << ((used & 3u) * 8u);
            node.qhi[a * 2 + (used >> 2)] |= qh << ((used & 3u) * 8u);
        }
        node.child[used++] = ref;
    }
    for (uint32_t k = used; k < (uint32_t)kBvhWidth; k++) node.child[k] = kEmptyRef;
    node.packed = e[0] | (e[1] << 8) | (e[2] << 16) | (used << 24);
    nodes[it.wide] = node;
}

__global__ __launch_bounds__(kB) void kReorderTris(const float* tris9, const uint32_t* sortedTri, uint32_t n,
                                                  TriRecord* out) {
    uint32_t i = blockIdx.x * kB + threadIdx.x;
    if (i >= n) return;
    uint32_t t = sortedTri[i];
    const float* p = tris9 + (size_t)t * 9u;
    TriRecord r;
    r.v0 = make_float4(p[0], p[1], p[2], __uint_as_float(t));
    r.v1 = make_float4(p[3], p[4], p[5], 0.0f);
    r.v2 = make_float4(p[6], p[7], p[8], 0.0f);
    out[i] = r;
}

__global__ void kInitCollapse(CollapseItem* items, uint32_t* counts, int n) {
    items[0] = CollapseItem{0, 0u};  // binary root: internal node 0, or the single leaf when n == 1
    counts[0] = 1u;  // items in the current frontier
    counts[1] = 0u;  // items in the next frontier
    counts[2] = 1u;  // wide nodes allocated
    counts[3] = 0u;  // overflow
}

__global__ void kFillU32(uint32_t* p, size_t n, uint32_t v) {
    size_t i = (size_t)blockIdx.x * kB + threadIdx.x;
    if (i < n) p[i] = v;
}

static uint32_t blocks(size_t n) { return (uint32_t)((n + kB - 1) / kB); }

}  // namespace mrt

using namespace mrt;

extern "C" int mrt_build_bvh(mrt_context* ctx, mrt_bvh_stats* stats) {
    if (!ctx) return -1;
    if (!check(hipSetDevice(ctx->device), "hipSetDevice")) return -2;
    hipStream_t s = ctx->stream;
    uint32_t n = ctx->triCount;
    if (stats) memset(stats, 0, sizeof(*stats));
    if (n == 0) {
        ctx->nodes.count = 0;
        ctx->tris.count = 0;
        ctx->bvhMaxDepth = 0;
        return 0;
    }
    hipEventRecord(ctx->evStage[0], s);
    DeviceArray<Aabb> triBoxes, nodeBoxes;
    DeviceArray<uint32_t> cb, vals, valsSorted, arrival, counts;
    DeviceArray<uint64_t> keys, keysSorted;
    DeviceArray<int2> children, range;
    DeviceArray<int> parent;
    DeviceArray<CollapseItem> front[2];
    int nn = (int)n;
    size_t binNodes = (size_t)2 * n - 1;
    size_t wideCap = n;  // every wide node but the root is a distinct binary node with > kLeafMaxTris triangles
    bool ok = triBoxes.reserve(n) && nodeBoxes.reserve(binNodes) && cb.reserve(6) && vals.reserve(n) &&
              valsSorted.reserve(n) && arrival.reserve(n) && counts.reserve(4) && keys.reserve(n) &&
              keysSorted.reserve(n) && children.reserve(n) && range.reserve(n) && parent.reserve(binNodes) &&
              front[0].reserve(wideCap) && front[1].reserve(wideCap) && ctx->nodes.reserve(wideCap) &&
              ctx->tris.reserve(n);
    auto cleanup = [&]() {
        triBoxes.release();
        nodeBoxes.release();
        cb.release();
        vals.release();
        valsSorted.release();
        arrival.release();
        counts.release();
        keys.release();
        keysSorted.release();
        children.release();
        range.release();
        parent.release();
        front[0].release();
        front[1].release();
    };
    if (!ok) {
        cleanup();
        return -3;
    }
    uint32_t initCb[6] = {0xffffffffu, 0xffffffffu, 0xffffffffu, 0u, 0u, 0u};
    hipMemcpyAsync(cb.ptr, initCb, sizeof(initCb), hipMemcpyHostToDevice, s);
    hipLaunchKernelGGL(kTriBounds, dim3(blocks(n)), dim3(kB), 0, s, ctx->worldTris.ptr, n, triBoxes.ptr, cb.ptr);
    hipLaunchKernelGGL(kMorton, dim3(blocks(n)), dim3(kB), 0, s, triBoxes.ptr, n, cb.ptr, keys.ptr, vals.ptr);
    size_t tempBytes = 0;
    rocprim::radix_sort_pairs(nullptr, tempBytes, keys.ptr, keysSorted.ptr, vals.ptr, valsSorted.ptr, n, 0, 63, s);
    void* temp = nullptr;
    if (!check(hipMalloc(&temp, tempBytes ? tempBytes : 1), "hipMalloc sort")) {
        cleanup();
        return -3;
    }
    rocprim::radix_sort_pairs(temp, tempBytes, keys.ptr, keysSorted.ptr, vals.ptr, valsSorted.ptr, n, 0, 63, s);
    hipLaunchKernelGGL(kFillU32, dim3(blocks(binNodes)), dim3(kB), 0, s, reinterpret_cast<uint32_t*>(parent.ptr),
                       binNodes, 0xffffffffu);
    if (n > 1)
        hipLaunchKernelGGL(kRadixTree, dim3(blocks(n - 1)), dim3(kB), 0, s, keysSorted.ptr, nn, children.ptr,
                           parent.ptr, range.ptr);
    hipLaunchKernelGGL(kFillU32, dim3(blocks(n)), dim3(kB), 0, s, arrival.ptr, (size_t)n, 0u);
    hipLaunchKernelGGL(kRefit, dim3(blocks(n)), dim3(kB), 0, s, triBoxes.ptr, valsSorted.ptr, nn, children.ptr,
                       parent.ptr, nodeBoxes.ptr, arrival.ptr);
    hipLaunchKernelGGL(kReorderTris, dim3(blocks(n)), dim3(kB), 0, s, ctx->worldTris.ptr, valsSorted.ptr, n,
                       ctx->tris.ptr);
    hipLaunchKernelGGL(kInitCollapse, dim3(1), dim3(1), 0, s, front[0].ptr, counts.ptr, nn);
    uint32_t depth = 0;
    int cur = 0;
    uint32_t hostCounts[4] = {1, 0, 1, 0};
    while (true) {
        uint32_t frontier = hostCounts[0];
        if (frontier == 0) break;
        depth++;
        hipMemsetAsync(counts.ptr + 1, 0, sizeof(uint32_t), s);
        hipLaunchKernelGGL(kCollapse, dim3(blocks(frontier)), dim3(kB), 0, s, front[cur].ptr, counts.ptr,
                           front[1 - cur].ptr, counts.ptr + 1, counts.ptr + 2, nn, children.ptr, range.ptr,
                           nodeBoxes.ptr, ctx->nodes.ptr, (uint32_t)wideCap, counts.ptr + 3);
        // the next frontier becomes the current one
        hipMemcpyAsync(counts.ptr, counts.ptr + 1, sizeof(uint32_t), hipMemcpyDeviceToDevice, s);
        if (!check(hipMemcpyAsync(hostCounts, counts.ptr, sizeof(hostCounts), hipMemcpyDeviceToHost, s), "collapse") ||
            !check(hipStreamSynchronize(s), "collapse sync")) {
            hipFree(temp);
            cleanup();
            return -4;
        }
        cur = 1 - cur;
        if (depth > 4096) break;
    }
    hipEventRecord(ctx->evStage[1], s);
    hipStreamSynchronize(s);
    hipFree(temp);
    if (hostCounts[3] != 0) {
        cleanup();
        setError("BVH collapse exceeded node capacity");
        return -5;
    }
    ctx->nodes.count = hostCounts[2];
    ctx->tris.count = n;
    ctx->bvhMaxDepth = depth;
    float ms = 0.0f;
    hipEventElapsedTime(&ms, ctx->evStage[0], ctx->evStage[1]);
    if (stats) {
        stats->buildMs = ms;
        stats->nodeCount = hostCounts[2];
        stats->leafTriangles = n;
        stats->maxDepth = depth;
    }
    cleanup();
    return check(hipGetLastError(), "bvh build") ? 0 : -4;
}
