// On-GPU BVH build for gfx950, replacing the driver-built BLAS/TLAS of the
// reference (src/core/render/accel/blas.c:222-262, tlas.c:535-559).
//
//  1. per-triangle bounds and the centroid bounds of the scene
//  2. 63-bit Morton codes (21 bits per axis) of triangle centroids
//  3. radix sort of (code, triangle) pairs (rocPRIM)
//  4. binary tree by parallel locally-ordered clustering over the Morton order
//     (PLOC, Meister and Bittner, TVCG 2018): every cluster finds the
//     neighbour within kPlocRadius positions whose merged box has the least
//     surface area, mutual nearest neighbours merge, survivors are compacted;
//     repeated until one cluster remains. This refines the LBVH order into a
//     tree of surface-area-heuristic quality.
//  5. top-down collapse into 8-wide nodes: a node is opened greedily by
//     splitting its largest-area child until eight children remain; subtrees
//     of at most leafMax triangles (kLeafMaxTris, or kLeafMaxTrisLarge for
//     scenes above kLargeScene triangles) become leaves whose triangles are
//     written contiguously (depth-first order) as 48-byte records; child
//     boxes are quantised to 8 bits per plane, conservatively
#include <cstring>
#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_select.hpp>

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

__device__ __forceinline__ float halfArea(const Aabb& b) {
    float dx = fmaxf(b.hi[0] - b.lo[0], 0.0f), dy = fmaxf(b.hi[1] - b.lo[1], 0.0f), dz = fmaxf(b.hi[2] - b.lo[2], 0.0f);
    return dx * dy + dy * dz + dz * dx;
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
// the exact planes origin + q * step (step a power of two, so q * step is
// exact and the sum is exact in double precision) satisfy lo <= clo and
// hi >= chi, with no rounding in between: the traversal folds the decode into
// its slab distances and never rounds a plane on its own.
__device__ void quantizePlanes(float origin, float step, float clo, float chi, uint32_t& qlo, uint32_t& qhi) {
    float a = floorf((clo - origin) / step), b = ceilf((chi - origin) / step);
    int l = (int)fminf(fmaxf(a, 0.0f), 255.0f), h = (int)fminf(fmaxf(b, 0.0f), 255.0f);
    while (l > 0 && (double)l * (double)step + (double)origin > (double)clo) l--;
    while (h < 255 && (double)h * (double)step + (double)origin < (double)chi) h++;
    qlo = (uint32_t)l;
    qhi = (uint32_t)h;
}

// ---------------------------------------------------------------- PLOC --
// Neighbourhood searched for each cluster's nearest neighbour. With one
// triangle per leaf, radii of 4, 8, 12 and 16 gave 7.97, 8.02, 8.12 and 8.39
// node visits per ray on the 1080p spectral workload (12.60, 12.57, 12.52
// and 13.01 on the 20.8M-triangle one) and 4 traced fastest on both.
constexpr int kPlocRadius = 4;

__device__ __forceinline__ Aabb merge(const Aabb& a, const Aabb& b) {
    Aabb m;
    for (int k = 0; k < 3; k++) {
        m.lo[k] = fminf(a.lo[k], b.lo[k]);
        m.hi[k] = fmaxf(a.hi[k], b.hi[k]);
    }
    return m;
}

// Leaf i is sorted triangle i; nodes n.. are created by merges.
__global__ __launch_bounds__(kB) void kPlocInit(const Aabb* triBoxes, const uint32_t* sortedTri, uint32_t n,
                                               Aabb* boxes, uint32_t* counts, uint32_t* clusters) {
    uint32_t i = blockIdx.x * kB + threadIdx.x;
    if (i >= n) return;
    boxes[i] = triBoxes[sortedTri[i]];
    counts[i] = 1u;
    clusters[i] = i;
}

// Pairs are ordered by (merged surface area, lower index, higher index), so the
// globally smallest pair is always a mutual nearest-neighbour pair and every
// pass merges at least once.
__device__ __forceinline__ bool pairBefore(float c0, int i, int j0, float c1, int j1) {
    if (c0 != c1) return c0 < c1;
    int a0 = min(i, j0), b0 = max(i, j0), a1 = min(i, j1), b1 = max(i, j1);
    return a0 != a1 ? a0 < a1 : b0 < b1;
}

__global__ __launch_bounds__(kB) void kPlocNearest(const uint32_t* clusters, uint32_t m, const Aabb* boxes, int* nn) {
    __shared__ Aabb tile[kB + 2 * kPlocRadius];
    const int b0 = (int)(blockIdx.x * kB);
    for (int t = threadIdx.x; t < kB + 2 * kPlocRadius; t += kB) {
        int g = b0 - kPlocRadius + t;
        if (g >= 0 && g < (int)m) tile[t] = boxes[clusters[g]];
    }
    __syncthreads();
    const int i = b0 + (int)threadIdx.x;
    if (i >= (int)m) return;
    const Aabb a = tile[threadIdx.x + kPlocRadius];
    float best = INFINITY;
    int bj = -1;
    for (int k = -kPlocRadius; k <= kPlocRadius; k++) {
        int j = i + k;
        if (k == 0 || j < 0 || j >= (int)m) continue;
        float c = halfArea(merge(a, tile[threadIdx.x + kPlocRadius + k]));
        if (bj < 0 || pairBefore(c, i, j, best, bj)) {
            best = c;
            bj = j;
        }
    }
    nn[i] = bj;
}

__global__ __launch_bounds__(kB) void kPlocMerge(uint32_t* clusters, uint32_t m, const int* nn, Aabb* boxes,
                                                uint32_t* counts, int2* children, uint32_t n, uint32_t* nodeCounter,
                                                uint8_t* keep) {
    const int i = (int)(blockIdx.x * kB + threadIdx.x);
    if (i >= (int)m) return;
    const int j = nn[i];
    const bool mutual = j >= 0 && nn[j] == i;
    if (mutual && i < j) {
        const uint32_t a = clusters[i], b = clusters[j];
        const uint32_t id = atomicAdd(nodeCounter, 1u);
        children[id - n] = make_int2((int)a, (int)b);
        boxes[id] = merge(boxes[a], boxes[b]);
        counts[id] = counts[a] + counts[b];
        clusters[i] = id;
    }
    keep[i] = (mutual && i > j) ? 0u : 1u;
}

// ------------------------------------------------------------- collapse --
struct CollapseItem {
    uint32_t node;   // binary node: leaf (< n) or merged node (>= n)
    uint32_t wide;   // wide node slot to fill
    uint32_t first;  // first triangle slot of the subtree in leaf order
    uint32_t pad;
};

__device__ __forceinline__ void writeTriangle(const float* tris9, uint32_t t, TriRecord* out) {
    const float* p = tris9 + (size_t)t * 9u;
    TriRecord r;
    r.v0 = make_float4(p[0], p[1], p[2], __uint_as_float(t));
    r.v1 = make_float4(p[3], p[4], p[5], 0.0f);
    r.v2 = make_float4(p[6], p[7], p[8], 0.0f);
    *out = r;
}

__global__ __launch_bounds__(kB) void kCollapse(const CollapseItem* in, const uint32_t* inCount, CollapseItem* out,
                                               uint32_t* outCount, uint32_t* nodeAlloc, uint32_t n,
                                               const int2* children, const uint32_t* counts, const Aabb* boxes,
                                               const uint32_t* sortedTri, const float* tris9, TriRecord* triOut,
                                               BvhNode* nodes, uint32_t nodeCapacity, uint32_t* overflow,
                                               uint32_t leafMax) {
    uint32_t w = blockIdx.x * kB + threadIdx.x;
    if (w >= *inCount) return;
    const CollapseItem it = in[w];
    uint32_t slot[kBvhWidth], first[kBvhWidth];
    int cnt = 0;
    if (it.node < n || counts[it.node] <= leafMax) {
        slot[cnt] = it.node;  // only the root of a tiny scene
        first[cnt++] = it.first;
    } else {
        int2 c = children[it.node - n];
        slot[0] = (uint32_t)c.x;
        first[0] = it.first;
        slot[1] = (uint32_t)c.y;
        first[1] = it.first + counts[c.x];
        cnt = 2;
        while (cnt < kBvhWidth) {
            // open the largest child that has too many triangles for a leaf
            int best = -1;
            float bestArea = -1.0f;
            for (int k = 0; k < cnt; k++) {
                if (counts[slot[k]] <= leafMax) continue;
                float a = halfArea(boxes[slot[k]]);
                if (a > bestArea) {
                    bestArea = a;
                    best = k;
                }
            }
            if (best < 0) break;
            int2 c2 = children[slot[best] - n];
            uint32_t f = first[best];
            slot[best] = (uint32_t)c2.x;
            slot[cnt] = (uint32_t)c2.y;
            first[cnt++] = f + counts[c2.x];
        }
    }
    const Aabb P = boxes[it.node];
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
        const uint32_t nd = slot[k];
        const uint32_t count = counts[nd];
        uint32_t ref;
        if (count <= leafMax) {
            // the subtree's triangles, depth first, into [first, first + count)
            uint32_t stack[kLeafMaxTrisLarge];
            int sp = 0;
            uint32_t written = 0;
            stack[sp++] = nd;
            while (sp > 0) {
                uint32_t x = stack[--sp];
                if (x < n) {
                    writeTriangle(tris9, sortedTri[x], &triOut[first[k] + written]);
                    written++;
                } else {
                    int2 c = children[x - n];
                    stack[sp++] = (uint32_t)c.y;
                    stack[sp++] = (uint32_t)c.x;
                }
            }
            ref = kLeafBit | ((count - 1u) << 27) | first[k];
        } else {
            uint32_t child = atomicAdd(nodeAlloc, 1u);
            if (child >= nodeCapacity) {
                atomicAdd(overflow, 1u);
                continue;
            }
            ref = child;
            uint32_t o = atomicAdd(outCount, 1u);
            out[o] = CollapseItem{nd, child, first[k], 0u};
        }
        const Aabb b = boxes[nd];
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

__global__ void kInitCollapse(CollapseItem* items, uint32_t* counts, const uint32_t* root) {
    items[0] = CollapseItem{*root, 0u, 0u, 0u};
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
    const size_t binNodes = (size_t)2 * n - 1;
    const size_t wideCap = n;  // every wide node but the root is a distinct binary node with more than one triangle
    DeviceArray<Aabb> triBoxes, boxes;
    DeviceArray<uint32_t> cb, vals, valsSorted, counts, nodeCounts, clusters[2], selected;
    DeviceArray<uint64_t> keys, keysSorted;
    DeviceArray<int2> children;
    DeviceArray<int> nn;
    DeviceArray<uint8_t> keep;
    DeviceArray<CollapseItem> front[2];
    bool ok = triBoxes.reserve(n) && boxes.reserve(binNodes) && cb.reserve(6) && vals.reserve(n) &&
              valsSorted.reserve(n) && counts.reserve(4) && nodeCounts.reserve(binNodes) && clusters[0].reserve(n) &&
              clusters[1].reserve(n) && selected.reserve(2) && keys.reserve(n) && keysSorted.reserve(n) &&
              children.reserve(n) && nn.reserve(n) && keep.reserve(n) && front[0].reserve(wideCap) &&
              front[1].reserve(wideCap) && ctx->nodes.reserve(wideCap) && ctx->tris.reserve(n);
    void* temp = nullptr;
    auto cleanup = [&]() {
        if (temp) hipFree(temp);
        temp = nullptr;
        triBoxes.release();
        boxes.release();
        cb.release();
        vals.release();
        valsSorted.release();
        counts.release();
        nodeCounts.release();
        clusters[0].release();
        clusters[1].release();
        selected.release();
        keys.release();
        keysSorted.release();
        children.release();
        nn.release();
        keep.release();
        front[0].release();
        front[1].release();
    };
    if (!ok) {
        cleanup();
        return -3;
    }
    // 1-3: bounds, Morton codes, sort
    uint32_t initCb[6] = {0xffffffffu, 0xffffffffu, 0xffffffffu, 0u, 0u, 0u};
    hipMemcpyAsync(cb.ptr, initCb, sizeof(initCb), hipMemcpyHostToDevice, s);
    hipLaunchKernelGGL(kTriBounds, dim3(blocks(n)), dim3(kB), 0, s, ctx->worldTris.ptr, n, triBoxes.ptr, cb.ptr);
    hipLaunchKernelGGL(kMorton, dim3(blocks(n)), dim3(kB), 0, s, triBoxes.ptr, n, cb.ptr, keys.ptr, vals.ptr);
    size_t sortBytes = 0, selectBytes = 0;
    rocprim::radix_sort_pairs(nullptr, sortBytes, keys.ptr, keysSorted.ptr, vals.ptr, valsSorted.ptr, n, 0, 63, s);
    rocprim::select(nullptr, selectBytes, clusters[0].ptr, keep.ptr, clusters[1].ptr, selected.ptr, (size_t)n, s);
    if (!check(hipMalloc(&temp, std::max<size_t>(std::max(sortBytes, selectBytes), 1)), "hipMalloc build scratch")) {
        cleanup();
        return -3;
    }
    rocprim::radix_sort_pairs(temp, sortBytes, keys.ptr, keysSorted.ptr, vals.ptr, valsSorted.ptr, n, 0, 63, s);

    // 4: PLOC
    hipLaunchKernelGGL(kPlocInit, dim3(blocks(n)), dim3(kB), 0, s, triBoxes.ptr, valsSorted.ptr, n, boxes.ptr,
                       nodeCounts.ptr, clusters[0].ptr);
    uint32_t nodeCounterInit = n;
    hipMemcpyAsync(counts.ptr + 3, &nodeCounterInit, sizeof(uint32_t), hipMemcpyHostToDevice, s);
    uint32_t m = n;
    int cur = 0;
    uint32_t passes = 0;
    while (m > 1) {
        hipLaunchKernelGGL(kPlocNearest, dim3(blocks(m)), dim3(kB), 0, s, clusters[cur].ptr, m, boxes.ptr, nn.ptr);
        hipLaunchKernelGGL(kPlocMerge, dim3(blocks(m)), dim3(kB), 0, s, clusters[cur].ptr, m, nn.ptr, boxes.ptr,
                           nodeCounts.ptr, children.ptr, n, counts.ptr + 3, keep.ptr);
        size_t bytes = selectBytes;
        rocprim::select(temp, bytes, clusters[cur].ptr, keep.ptr, clusters[1 - cur].ptr, selected.ptr, (size_t)m, s);
        uint32_t next = 0;
        if (!check(hipMemcpyAsync(&next, selected.ptr, sizeof(uint32_t), hipMemcpyDeviceToHost, s), "ploc count") ||
            !check(hipStreamSynchronize(s), "ploc sync")) {
            cleanup();
            return -4;
        }
        cur = 1 - cur;
        passes++;
        if (next >= m || passes > 100000u) {  // every pass merges at least one pair
            cleanup();
            setError("BVH clustering made no progress");
            return -5;
        }
        m = next;
    }

    // the root cluster's box, padded outward by 1e-5 of its size and coordinates
    // so that rounding in the rays' slab test can never reject a ray that
    // reaches a triangle
    {
        uint32_t root = 0;
        Aabb rb{};
        if (!check(hipMemcpyAsync(&root, clusters[cur].ptr, sizeof(root), hipMemcpyDeviceToHost, s), "root") ||
            !check(hipStreamSynchronize(s), "root sync") ||
            !check(hipMemcpy(&rb, boxes.ptr + root, sizeof(rb), hipMemcpyDeviceToHost), "root box")) {
            cleanup();
            return -4;
        }
        for (int a = 0; a < 3; a++) {
            const float ext = rb.hi[a] - rb.lo[a];
            const bool finite = std::isfinite(rb.lo[a]) && std::isfinite(rb.hi[a]) && ext >= 0.0f;
            ctx->sceneBoxLo[a] = finite ? rb.lo[a] - 1e-5f * (fabsf(rb.lo[a]) + ext) - 1e-30f : -INFINITY;
            ctx->sceneBoxHi[a] = finite ? rb.hi[a] + 1e-5f * (fabsf(rb.hi[a]) + ext) + 1e-30f : INFINITY;
        }
    }

    // 5: collapse into 8-wide nodes, writing triangle records in leaf order
    hipLaunchKernelGGL(kInitCollapse, dim3(1), dim3(1), 0, s, front[0].ptr, counts.ptr, clusters[cur].ptr);
    uint32_t depth = 0;
    int fc = 0;
    uint32_t hostCounts[4] = {1, 0, 1, 0};
    while (true) {
        uint32_t frontier = hostCounts[0];
        if (frontier == 0) break;
        depth++;
        hipMemsetAsync(counts.ptr + 1, 0, sizeof(uint32_t), s);
        hipLaunchKernelGGL(kCollapse, dim3(blocks(frontier)), dim3(kB), 0, s, front[fc].ptr, counts.ptr,
                           front[1 - fc].ptr, counts.ptr + 1, counts.ptr + 2, n, children.ptr, nodeCounts.ptr,
                           boxes.ptr, valsSorted.ptr, ctx->worldTris.ptr, ctx->tris.ptr, ctx->nodes.ptr,
                           (uint32_t)wideCap, counts.ptr + 3, n > kLargeScene ? kLeafMaxTrisLarge : kLeafMaxTris);
        // the next frontier becomes the current one
        hipMemcpyAsync(counts.ptr, counts.ptr + 1, sizeof(uint32_t), hipMemcpyDeviceToDevice, s);
        if (!check(hipMemcpyAsync(hostCounts, counts.ptr, sizeof(hostCounts), hipMemcpyDeviceToHost, s), "collapse") ||
            !check(hipStreamSynchronize(s), "collapse sync")) {
            cleanup();
            return -4;
        }
        fc = 1 - fc;
        if (depth > 4096) break;
    }
    hipEventRecord(ctx->evStage[1], s);
    hipStreamSynchronize(s);
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
