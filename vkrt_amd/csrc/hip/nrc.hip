// Neural radiance cache for gfx950: a fully connected network with FP16 weights
// on MFMA (v_mfma_f32_16x16x32_f16 / 16x16x16_f16) and FP32 accumulation,
// predicting the radiance a surface point scatters towards a direction
// (Mueller, Rousselle, Novak and Keller, "Real-time Neural Radiance Caching for
// Path Tracing", SIGGRAPH 2021). This is an addition: the reference has no
// neural cache.
//
// Input: 11 values in [0, 1] per sample -- position (3, scene box), outgoing
// direction (2, theta/pi and phi/2pi), shading normal (2, same), roughness,
// albedo (3). The first seven are encoded with 12 frequencies,
// sin(2^f pi v) and cos(2^f pi v) (168 features), followed by roughness and
// albedo (4), a constant 1 that acts as the first layer's bias, and zero
// padding to 192. Network: 192 -> 64 -> 64 -> 64 -> 64 -> 64 -> 3, ReLU on the
// hidden layers, no biases, linear output.
//
// Fragment layouts (checked on the device): for 16x16x32, lane l holds
// A[l%16][8(l/16)+t] and B[8(l/16)+t][l%16] (t < 8); for 16x16x16, A[l%16][4(l/16)+t]
// and B[4(l/16)+t][l%16] (t < 4); both produce D[4(l/16)+t][l%16] (t < 4). A layer's
// D fragments become the next layer's B fragments without moving data across
// lanes: the 8 K-slots of lane group g in K-block kb are taken to be hidden
// neurons 32kb + 4g + {0..3} and 32kb + 16 + 4g + {0..3} (two consecutive D
// blocks), and the weight columns are stored in that order.
#include <algorithm>
#include <cmath>

#include "dev_shading.h"
#include "mrt_internal.hpp"

namespace mrt {

typedef _Float16 half4 __attribute__((ext_vector_type(4)));
typedef _Float16 half8 __attribute__((ext_vector_type(8)));
typedef float float4v __attribute__((ext_vector_type(4)));

constexpr int kNrcIn = 192;      // encoded input width
constexpr int kNrcWidth = 64;    // hidden width
constexpr int kNrcHidden = 5;    // hidden layers
constexpr int kNrcOutRows = 16;  // 3 outputs padded to one MFMA block
constexpr int kNrcFreqs = 12;
// natural fp32 parameter layout: W0 [64][192], W1..W4 [64][64], W5 [16][64] (rows 3..15 zero)
constexpr int kNrcOffW0 = 0;
constexpr int kNrcOffHidden = kNrcWidth * kNrcIn;  // 12288
constexpr int kNrcOffOut = kNrcOffHidden + 4 * kNrcWidth * kNrcWidth;  // 28672
constexpr int kNrcParams = kNrcOffOut + kNrcOutRows * kNrcWidth;  // 29696
// transposed fp16 copies for back-propagation: W1..W4 as [in][out slot], W5 as [in][16]
constexpr int kNrcBwdOut = 4 * kNrcWidth * kNrcWidth;  // 16384
constexpr int kNrcBwdParams = kNrcBwdOut + kNrcWidth * kNrcOutRows;  // 17408
constexpr int kNrcBlock = 256;
// LDS copies of the fp16 weights, rows padded by 16 bytes so that the 16 rows a
// fragment read touches start 4 banks apart (the dense rows of 384 and 128
// bytes would put them all on the same few banks): forward W0 rows of 200
// halfs, W1..W4 rows of 72, W5 dense; backward W1^T..W4^T rows of 72, W5^T rows of 20.
constexpr int kNrcLdsStrideIn = kNrcIn + 8, kNrcLdsStrideW = kNrcWidth + 8, kNrcLdsStrideOut = kNrcOutRows + 4;
constexpr int kNrcLdsOffHidden = kNrcWidth * kNrcLdsStrideIn;                      // 12800
constexpr int kNrcLdsOffOut = kNrcLdsOffHidden + 4 * kNrcWidth * kNrcLdsStrideW;   // 31232
constexpr int kNrcLdsFwd = kNrcLdsOffOut + kNrcOutRows * kNrcWidth;                // 32256
constexpr int kNrcLdsBwdOut = 4 * kNrcWidth * kNrcLdsStrideW;                      // 18432
constexpr int kNrcLdsBwd = kNrcLdsBwdOut + kNrcWidth * kNrcLdsStrideOut;           // 19712

// hidden neuron held in K-slot s of a 64-wide hidden input (see the header)
__host__ __device__ inline int nrcNeuronOfSlot(int s) {
    int kb = s >> 5, r = s & 31, g = r >> 3, t = r & 7;
    return 32 * kb + (t < 4 ? 4 * g + t : 16 + 4 * g + (t - 4));
}
__host__ __device__ inline int nrcSlotOfNeuron(int n) {
    int kb = n >> 5, r = n & 31;
    int g, t;
    if (r < 16) {
        g = r >> 2;
        t = r & 3;
    } else {
        g = (r - 16) >> 2;
        t = 4 + ((r - 16) & 3);
    }
    return 32 * kb + 8 * g + t;
}

__device__ inline void nrcLoadRaw(const float4* in, uint32_t stride, uint32_t idx, bool valid, float* v) {
    if (!valid) {
        for (int k = 0; k < 11; k++) v[k] = 0.0f;
        return;
    }
    const float4* r = in + (size_t)stride * idx;
    float4 a = r[0], b = r[1], c = r[2];
    v[0] = a.x; v[1] = a.y; v[2] = a.z; v[3] = a.w;
    v[4] = b.x; v[5] = b.y; v[6] = b.z; v[7] = b.w;
    v[8] = c.x; v[9] = c.y; v[10] = c.z;
}

// v[d] for a lane-dependent d < 7 without indexing a register array
__device__ inline float nrcPick(const float* v, int d) {
    float r = v[0];
    r = d == 1 ? v[1] : r;
    r = d == 2 ? v[2] : r;
    r = d == 3 ? v[3] : r;
    r = d == 4 ? v[4] : r;
    r = d == 5 ? v[5] : r;
    r = d == 6 ? v[6] : r;
    return r;
}

// B fragments of the encoded input of sample l%16: K-block kb holds features
// 32kb + 8g + t. Feature 24d + 2f (+1) is sin (cos) of 2^f pi v[d], computed as
// the hardware sine / cosine of frac(2^(f-1) v[d]) revolutions (the scaling is
// exact); 168..171 are roughness and albedo, 172 the constant 1, the rest 0.
__device__ inline void nrcEncode(const float* v, int g, half8 x[6]) {
#pragma unroll
    for (int kb = 0; kb < 6; kb++) {
        const int f0 = 32 * kb + 8 * g;
        half8 h = {0, 0, 0, 0, 0, 0, 0, 0};
        if (f0 < 7 * 2 * kNrcFreqs) {
            const float vd = nrcPick(v, f0 / (2 * kNrcFreqs));
            const int freq0 = (f0 % (2 * kNrcFreqs)) >> 1;
#pragma unroll
            for (int t = 0; t < 4; t++) {
                float u = ldexpf(vd, freq0 + t - 1);
                u -= floorf(u);
                h[2 * t] = (_Float16)__builtin_amdgcn_sinf(u);
                h[2 * t + 1] = (_Float16)__builtin_amdgcn_cosf(u);
            }
        } else if (f0 == 7 * 2 * kNrcFreqs) {
            h = half8{(_Float16)v[7], (_Float16)v[8], (_Float16)v[9], (_Float16)v[10], (_Float16)1.0f, 0, 0, 0};
        }
        x[kb] = h;
    }
}

// A zero the compiler cannot see through. Added to the weight pointers inside a
// sample loop it keeps the weight fragment loads in the loop instead of hoisted
// out of it, where all 58 fragments would occupy 232 registers.
__device__ inline int nrcOpaqueZero() {
    int z;
    asm volatile("s_mov_b32 %0, 0" : "=s"(z));
    return z;
}

__device__ inline half8 ldHalf8(const _Float16* p) { return *reinterpret_cast<const half8*>(p); }
__device__ inline half4 ldHalf4(const _Float16* p) { return *reinterpret_cast<const half4*>(p); }

__device__ inline float4v relu4(float4v a) {
    return float4v{fmaxf(a[0], 0.0f), fmaxf(a[1], 0.0f), fmaxf(a[2], 0.0f), fmaxf(a[3], 0.0f)};
}
// two consecutive D blocks (neurons 32kb + 4g + t and 32kb + 16 + 4g + t) as one K=32 B fragment
__device__ inline half8 nrcPack(float4v lo, float4v hi) {
    return half8{(_Float16)lo[0], (_Float16)lo[1], (_Float16)lo[2], (_Float16)lo[3],
                 (_Float16)hi[0], (_Float16)hi[1], (_Float16)hi[2], (_Float16)hi[3]};
}

// Copies of the forward (and backward) fp16 weights from their dense global
// layout into the padded LDS layout, by the block's threads.
__device__ inline void nrcLoadForward(_Float16* w, const _Float16* src, int nthreads) {
    for (int k = threadIdx.x; k < kNrcWidth * (kNrcIn / 8); k += nthreads) {
        const int r = k / (kNrcIn / 8), c = k % (kNrcIn / 8);
        *reinterpret_cast<half8*>(w + r * kNrcLdsStrideIn + 8 * c) =
            *reinterpret_cast<const half8*>(src + kNrcOffW0 + r * kNrcIn + 8 * c);
    }
    for (int k = threadIdx.x; k < 4 * kNrcWidth * (kNrcWidth / 8); k += nthreads) {
        const int r = k / (kNrcWidth / 8), c = k % (kNrcWidth / 8);
        *reinterpret_cast<half8*>(w + kNrcLdsOffHidden + r * kNrcLdsStrideW + 8 * c) =
            *reinterpret_cast<const half8*>(src + kNrcOffHidden + r * kNrcWidth + 8 * c);
    }
    for (int k = threadIdx.x; k < kNrcOutRows * kNrcWidth / 8; k += nthreads)
        reinterpret_cast<half8*>(w + kNrcLdsOffOut)[k] = reinterpret_cast<const half8*>(src + kNrcOffOut)[k];
}
__device__ inline void nrcLoadBackward(_Float16* w, const _Float16* src, int nthreads) {
    for (int k = threadIdx.x; k < 4 * kNrcWidth * (kNrcWidth / 8); k += nthreads) {
        const int r = k / (kNrcWidth / 8), c = k % (kNrcWidth / 8);
        *reinterpret_cast<half8*>(w + r * kNrcLdsStrideW + 8 * c) = *reinterpret_cast<const half8*>(src + r * kNrcWidth + 8 * c);
    }
    for (int k = threadIdx.x; k < kNrcWidth * (kNrcOutRows / 4); k += nthreads) {
        const int r = k / (kNrcOutRows / 4), c = k % (kNrcOutRows / 4);
        *reinterpret_cast<half4*>(w + kNrcLdsBwdOut + r * kNrcLdsStrideOut + 4 * c) =
            *reinterpret_cast<const half4*>(src + kNrcBwdOut + r * kNrcOutRows + 4 * c);
    }
}

// Forward pass of 16 samples (one per lane%16). w: the padded LDS copy of the forward weights.
// Saves the hidden activations as B fragments when h != nullptr. Returns the output D fragment.
template <bool SAVE>
__device__ inline float4v nrcForward(const _Float16* w, const half8 x[6], half8 (*h)[2]) {
    const int l = threadIdx.x & 63, row = l & 15, g = l >> 4;
    float4v d[4];
#pragma unroll
    for (int ib = 0; ib < 4; ib++) {
        float4v acc = {0.0f, 0.0f, 0.0f, 0.0f};
        const _Float16* wr = w + (ib * 16 + row) * kNrcLdsStrideIn + 8 * g;
#pragma unroll
        for (int kb = 0; kb < 6; kb++) acc = __builtin_amdgcn_mfma_f32_16x16x32_f16(ldHalf8(wr + 32 * kb), x[kb], acc, 0, 0, 0);
        d[ib] = relu4(acc);
    }
    half8 in[2] = {nrcPack(d[0], d[1]), nrcPack(d[2], d[3])};
    if (SAVE) {
        h[0][0] = in[0];
        h[0][1] = in[1];
    }
#pragma unroll
    for (int layer = 1; layer < kNrcHidden; layer++) {
        const _Float16* wl = w + kNrcLdsOffHidden + (layer - 1) * kNrcWidth * kNrcLdsStrideW;
#pragma unroll
        for (int ib = 0; ib < 4; ib++) {
            float4v acc = {0.0f, 0.0f, 0.0f, 0.0f};
            const _Float16* wr = wl + (ib * 16 + row) * kNrcLdsStrideW + 8 * g;
            acc = __builtin_amdgcn_mfma_f32_16x16x32_f16(ldHalf8(wr), in[0], acc, 0, 0, 0);
            acc = __builtin_amdgcn_mfma_f32_16x16x32_f16(ldHalf8(wr + 32), in[1], acc, 0, 0, 0);
            d[ib] = relu4(acc);
        }
        in[0] = nrcPack(d[0], d[1]);
        in[1] = nrcPack(d[2], d[3]);
        if (SAVE) {
            h[layer][0] = in[0];
            h[layer][1] = in[1];
        }
    }
    float4v out = {0.0f, 0.0f, 0.0f, 0.0f};
    const _Float16* wo = w + kNrcLdsOffOut + row * kNrcWidth + 8 * g;
    out = __builtin_amdgcn_mfma_f32_16x16x32_f16(ldHalf8(wo), in[0], out, 0, 0, 0);
    out = __builtin_amdgcn_mfma_f32_16x16x32_f16(ldHalf8(wo + 32), in[1], out, 0, 0, 0);
    return out;  // lanes 0..15 hold outputs 0..3 of sample l (rows 0..2 used)
}

// Inference over n samples (n = min(*count, nFixed) when count is given).
// in: `stride` float4 per sample, the first three the raw inputs (in[2].w = path
// index bits when rad != null), in[3] the throughput when rad != null. Writes
// raw outputs to out (when non-null) and/or adds throughput * max(output, 0) to
// rad[path].
// rad != null: every query adds throughput x prediction to its path's radiance.
// spectral: the prediction is CIE XYZ; it is converted to linear sRGB, clipped
// at 0 and uplifted with the rgb2spec table at the query's four hero
// wavelengths (query word 4); word 3 holds the per-lane weights.
__global__ __launch_bounds__(kNrcBlock) void kNrcInfer(const _Float16* __restrict__ wfwd, const float4* __restrict__ in,
                                                      uint32_t stride, const uint32_t* count, uint32_t nFixed,
                                                      float4* out, float4* rad, Spec spec, uint32_t spectral) {
    __shared__ __align__(16) _Float16 w[kNrcLdsFwd];
    nrcLoadForward(w, wfwd, kNrcBlock);
    __syncthreads();
    const uint32_t n = count ? min(*count, nFixed) : nFixed;
    const int l = threadIdx.x & 63, g = l >> 4;
    const uint32_t wave = blockIdx.x * (kNrcBlock / 64) + (threadIdx.x >> 6);
    const uint32_t waves = gridDim.x * (kNrcBlock / 64);
    for (uint32_t base = wave * 16u; base < n; base += waves * 16u) {
        const uint32_t idx = base + (uint32_t)(l & 15);
        const bool valid = idx < n;
        float v[11];
        nrcLoadRaw(in, stride, idx, valid, v);
        half8 x[6];
        nrcEncode(v, g, x);
        float4v o = nrcForward<false>(w + nrcOpaqueZero(), x, nullptr);
        if (g == 0 && valid) {
            if (out) out[idx] = make_float4(o[0], o[1], o[2], 0.0f);
            if (rad) {
                const float4* rec = in + (size_t)stride * idx;
                const uint32_t path = __float_as_uint(rec[2].w);
                const float4 b = rec[3];
                float4 acc = rad[path];
                if (spectral) {
                    const F3 c = xyzToSrgb(F3{fmaxf(o[0], 0.0f), fmaxf(o[1], 0.0f), fmaxf(o[2], 0.0f)});
                    const F4 sp = spec4(spec, F3{fmaxf(c.x, 0.0f), fmaxf(c.y, 0.0f), fmaxf(c.z, 0.0f)}, ld4(rec[4]));
                    acc.x += b.x * sp.x;
                    acc.y += b.y * sp.y;
                    acc.z += b.z * sp.z;
                    acc.w += b.w * sp.w;
                } else {
                    acc.x += b.x * fmaxf(o[0], 0.0f);
                    acc.y += b.y * fmaxf(o[1], 0.0f);
                    acc.z += b.z * fmaxf(o[2], 0.0f);
                }
                rad[path] = acc;
            }
        }
    }
}

// Training runs in two passes over a batch.
//
// kNrcBackprop: every wave takes 16 samples, runs the forward pass keeping the
// activations, back-propagates the relative L2 loss sum_c (y_c - t_c)^2 /
// (y_c^2 + 0.01) (y detached in the denominator) and writes the fp16 layer
// inputs and deltas of its samples into feature-major tables (one row of `ld`
// halfs per feature, samples along the row):
//   rows   0..191  encoded input x          rows 512..831  deltas d1..d5 (64 each)
//   rows 192..511  activations h1..h5       rows 832..835  output deltas (3 used)
// kNrcWgrad: the weight gradients dW = sum over samples of delta x input^T are
// then plain GEMMs with the batch as the K dimension. Each wave takes one
// output strip of one layer over one chunk of kNrcChunk samples and writes its
// part of that chunk's partial gradient (grad + chunk * params); the optimiser
// step sums the chunks' partials.
constexpr int kNrcRowX = 0, kNrcRowH = kNrcIn, kNrcRowD = kNrcRowH + kNrcHidden * kNrcWidth;
constexpr int kNrcRowOut = kNrcRowD + kNrcHidden * kNrcWidth;  // 832
constexpr int kNrcTabRows = kNrcRowOut + 4;                    // 836
constexpr uint32_t kNrcChunk = 512;                            // samples per partial gradient
constexpr uint32_t kNrcSlice = 65536;                          // samples per pair of passes
constexpr int kNrcWgradTasks = 9;  // W0 in four 48-column strips, W1..W4, W5

__device__ inline void nrcStoreRows(_Float16* tab, uint32_t ld, int row0, uint32_t s, float4v v) {
#pragma unroll
    for (int t = 0; t < 4; t++) tab[(size_t)(row0 + t) * ld + s] = (_Float16)v[t];
}

// Both weight copies sit in LDS (104 KB, one block of eight waves per CU), and
// every wave takes one group of 16 samples.
constexpr int kNrcTrainBlock = 512;
constexpr size_t kNrcTrainLds = (size_t)(kNrcLdsFwd + kNrcLdsBwd) * sizeof(_Float16);

__global__ __launch_bounds__(kNrcTrainBlock) void kNrcBackprop(const _Float16* __restrict__ wfwd,
                                                              const _Float16* __restrict__ wbwd,
                                                              const float4* __restrict__ in, uint32_t stride,
                                                              const float4* __restrict__ target, const uint32_t* count,
                                                              uint32_t nFixed, uint32_t batch, uint32_t sliceBegin,
                                                              uint32_t batchOffset, uint32_t shuffle, _Float16* tab,
                                                              uint32_t ld, float* lossSum) {
    extern __shared__ __align__(16) _Float16 nrcW[];
    nrcLoadForward(nrcW, wfwd, kNrcTrainBlock);
    nrcLoadBackward(nrcW + kNrcLdsFwd, wbwd, kNrcTrainBlock);
    __syncthreads();
    const uint32_t n = count ? min(*count, nFixed) : nFixed;
    const int l = threadIdx.x & 63, row = l & 15, g = l >> 4;
    const uint32_t group = blockIdx.x * (kNrcTrainBlock / 64) + (threadIdx.x >> 6);
    float lossAcc = 0.0f;
    {
        const uint32_t s = group * 16u + (uint32_t)row;  // table column
        const uint32_t bi = sliceBegin + s;
        const bool valid = bi < batch && n > 0;
        // a pseudo-random training record for this batch slot, or record bi itself
        const uint32_t idx =
            !valid ? 0u : (shuffle ? (uint32_t)(((uint64_t)(batchOffset + bi) * 2654435761ull) % n) : bi % n);
        float v[11];
        nrcLoadRaw(in, stride, idx, valid, v);
        half8 x[6];
        nrcEncode(v, g, x);
#pragma unroll
        for (int kb = 0; kb < 6; kb++)
#pragma unroll
            for (int t = 0; t < 8; t++) tab[(size_t)(kNrcRowX + 32 * kb + 8 * g + t) * ld + s] = x[kb][t];
        half8 h[kNrcHidden][2];
        const int z = nrcOpaqueZero();
        float4v o = nrcForward<true>(nrcW + z, x, h);
#pragma unroll
        for (int layer = 0; layer < kNrcHidden; layer++)
#pragma unroll
            for (int kb = 0; kb < 2; kb++)
#pragma unroll
                for (int t = 0; t < 8; t++)
                    tab[(size_t)(kNrcRowH + layer * kNrcWidth + nrcNeuronOfSlot(32 * kb + 8 * g + t)) * ld + s] =
                        h[layer][kb][t];

        // loss gradient, D layout (rows 0..2 of lane group 0); invalid slots get zero deltas
        float4v dOut = {0.0f, 0.0f, 0.0f, 0.0f};
        const float4 tg = valid ? target[idx] : make_float4(0.0f, 0.0f, 0.0f, -1.0f);
        if (g == 0 && tg.w >= 0.0f) {
            float tv[3] = {tg.x, tg.y, tg.z};
#pragma unroll
            for (int c = 0; c < 3; c++) {
                float y = o[c], den = y * y + 0.01f, diff = y - tv[c];
                lossAcc += diff * diff / den;
                dOut[c] = fminf(fmaxf(2.0f * diff / den, -1000.0f), 1000.0f);
            }
        }
        const half4 dOutH = half4{(_Float16)dOut[0], (_Float16)dOut[1], (_Float16)dOut[2], (_Float16)dOut[3]};
        if (g == 0) nrcStoreRows(tab, ld, kNrcRowOut, s, dOut);

        // delta of h5: (W5^T dOut) masked by h5 > 0, D layout in natural neuron blocks
        const _Float16* wb = nrcW + kNrcLdsFwd + z;
        float4v dl[4];
#pragma unroll
        for (int mb = 0; mb < 4; mb++) {
            half4 a = ldHalf4(wb + kNrcLdsBwdOut + (mb * 16 + row) * kNrcLdsStrideOut + 4 * g);
            float4v acc = __builtin_amdgcn_mfma_f32_16x16x16f16(a, dOutH, float4v{0.0f, 0.0f, 0.0f, 0.0f}, 0, 0, 0);
#pragma unroll
            for (int t = 0; t < 4; t++) acc[t] = (float)h[4][mb >> 1][(mb & 1) * 4 + t] > 0.0f ? acc[t] : 0.0f;
            dl[mb] = acc;
        }
        // delta_l = (W_l^T delta_{l+1}) * (h_l > 0), layers 4..1
#pragma unroll
        for (int layer = kNrcHidden - 1; layer >= 1; layer--) {
#pragma unroll
            for (int mb = 0; mb < 4; mb++)
                nrcStoreRows(tab, ld, kNrcRowD + layer * kNrcWidth + mb * 16 + 4 * g, s, dl[mb]);
            const half8 bd[2] = {nrcPack(dl[0], dl[1]), nrcPack(dl[2], dl[3])};
            const _Float16* wbl = wb + (layer - 1) * kNrcWidth * kNrcLdsStrideW;
#pragma unroll
            for (int mb = 0; mb < 4; mb++) {
                const _Float16* wr = wbl + (mb * 16 + row) * kNrcLdsStrideW + 8 * g;
                float4v acc = {0.0f, 0.0f, 0.0f, 0.0f};
                acc = __builtin_amdgcn_mfma_f32_16x16x32_f16(ldHalf8(wr), bd[0], acc, 0, 0, 0);
                acc = __builtin_amdgcn_mfma_f32_16x16x32_f16(ldHalf8(wr + 32), bd[1], acc, 0, 0, 0);
#pragma unroll
                for (int t = 0; t < 4; t++)
                    acc[t] = (float)h[layer - 1][mb >> 1][(mb & 1) * 4 + t] > 0.0f ? acc[t] : 0.0f;
                dl[mb] = acc;
            }
        }
#pragma unroll
        for (int mb = 0; mb < 4; mb++) nrcStoreRows(tab, ld, kNrcRowD + mb * 16 + 4 * g, s, dl[mb]);
    }
    // one loss sum per wave
    for (int off = 32; off > 0; off >>= 1) lossAcc += __shfl_down(lossAcc, off);
    if (l == 0 && lossAcc != 0.0f) atomicAdd(lossSum, lossAcc);
}

// One strip of a weight gradient over one chunk: MB x NB tiles of 16 x 16,
// A = delta rows arow0.. (output neurons), B = input rows brow0.. (input
// features), K = the chunk's samples in steps of 32. ROWS3: the output layer,
// whose delta table has three live rows (the A rows of lanes 3..15 are zero).
template <int MB, int NB, bool ROWS3>
__device__ inline void nrcWgradStrip(const _Float16* __restrict__ tab, uint32_t ld, uint32_t k0, int arow0, int brow0,
                                     float* __restrict__ dst, int cols, int col0, bool accumulate) {
    const int l = threadIdx.x & 63, row = l & 15, g = l >> 4;
    float4v acc[MB][NB];
#pragma unroll
    for (int i = 0; i < MB; i++)
#pragma unroll
        for (int j = 0; j < NB; j++) acc[i][j] = float4v{0.0f, 0.0f, 0.0f, 0.0f};
    const bool liveA = !ROWS3 || row < 3;
    const _Float16* pa = tab + (size_t)(arow0 + (liveA ? row : 0)) * ld + k0 + 8 * g;
    const _Float16* pb = tab + (size_t)(brow0 + row) * ld + k0 + 8 * g;
#pragma unroll 2
    for (uint32_t k = 0; k < kNrcChunk; k += 32) {
        half8 a[MB], b[NB];
#pragma unroll
        for (int i = 0; i < MB; i++) a[i] = liveA ? ldHalf8(pa + (size_t)(16 * i) * ld + k) : half8{0, 0, 0, 0, 0, 0, 0, 0};
#pragma unroll
        for (int j = 0; j < NB; j++) b[j] = ldHalf8(pb + (size_t)(16 * j) * ld + k);
#pragma unroll
        for (int i = 0; i < MB; i++)
#pragma unroll
            for (int j = 0; j < NB; j++) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_f16(a[i], b[j], acc[i][j], 0, 0, 0);
    }
    // D[4g + t][row]: weight row 16i + 4g + t, column col0 + 16j + row
#pragma unroll
    for (int i = 0; i < MB; i++)
#pragma unroll
        for (int j = 0; j < NB; j++)
#pragma unroll
            for (int t = 0; t < 4; t++) {
                float* p = dst + (size_t)(16 * i + 4 * g + t) * cols + col0 + 16 * j + row;
                *p = accumulate ? *p + acc[i][j][t] : acc[i][j][t];
            }
}

__global__ __launch_bounds__(kNrcBlock) void kNrcWgrad(const _Float16* __restrict__ tab, uint32_t ld, uint32_t chunks,
                                                      float* grad, uint32_t accumulate) {
    const uint32_t wave = blockIdx.x * (kNrcBlock / 64) + (threadIdx.x >> 6);
    const uint32_t chunk = wave / kNrcWgradTasks, task = wave % kNrcWgradTasks;
    if (chunk >= chunks) return;
    float* part = grad + (size_t)chunk * kNrcParams;
    const uint32_t k0 = chunk * kNrcChunk;
    if (task < 4) {  // W0 [64][192], columns 48 task .. 48 task + 47
        nrcWgradStrip<4, 3, false>(tab, ld, k0, kNrcRowD, kNrcRowX + 48 * (int)task, part + kNrcOffW0, kNrcIn,
                                   48 * (int)task, accumulate != 0u);
    } else if (task < 8) {  // W_l [64][64] = d_{l+1} x h_l^T, l = 1..4
        const int layer = (int)task - 3;
        nrcWgradStrip<4, 4, false>(tab, ld, k0, kNrcRowD + layer * kNrcWidth, kNrcRowH + (layer - 1) * kNrcWidth,
                                   part + kNrcOffHidden + (layer - 1) * kNrcWidth * kNrcWidth, kNrcWidth, 0,
                                   accumulate != 0u);
    } else {  // W5 [16][64] = dOut x h5^T (rows 3..15 zero)
        nrcWgradStrip<1, 4, true>(tab, ld, k0, kNrcRowOut, kNrcRowH + (kNrcHidden - 1) * kNrcWidth,
                                  part + kNrcOffOut, kNrcWidth, 0, accumulate != 0u);
    }
}

// Sum over the chunks' partial gradients of parameter p, with 8
// independent loads in flight per step (a chain of single loads would wait on
// each one in turn).
__device__ inline float nrcSumParts(const float* grad, uint32_t parts, int p) {
    float acc[8] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
    uint32_t b = 0;
    for (; b + 8 <= parts; b += 8) {
#pragma unroll
        for (int u = 0; u < 8; u++) acc[u] += grad[(size_t)(b + u) * kNrcParams + p];
    }
    for (; b < parts; b++) acc[0] += grad[(size_t)b * kNrcParams + p];
    return ((acc[0] + acc[1]) + (acc[2] + acc[3])) + ((acc[4] + acc[5]) + (acc[6] + acc[7]));
}

// Sum of the chunks' partial gradients.
__global__ __launch_bounds__(kNrcBlock) void kNrcSumParts(const float* grad, uint32_t parts, float* out) {
    const int p = blockIdx.x * kNrcBlock + threadIdx.x;
    if (p >= kNrcParams) return;
    out[p] = nrcSumParts(grad, parts, p);
}

// Adam on the fp32 master weights (gradient = sum of `parts` partials times
// scale), then the fp16 forward and backward copies.
__global__ __launch_bounds__(kNrcBlock) void kNrcAdam(float* master, float* m1, float* m2, const float* grad,
                                                     uint32_t parts, _Float16* wfwd, _Float16* wbwd, float scale,
                                                     float lr, float b1, float b2, float c1, float c2, float eps) {
    const int p = blockIdx.x * kNrcBlock + threadIdx.x;
    if (p >= kNrcParams) return;
    int layer, i, k;
    if (p < kNrcOffHidden) {
        layer = 0;
        i = p / kNrcIn;
        k = p % kNrcIn;
    } else if (p < kNrcOffOut) {
        layer = 1 + (p - kNrcOffHidden) / (kNrcWidth * kNrcWidth);
        int r = (p - kNrcOffHidden) % (kNrcWidth * kNrcWidth);
        i = r / kNrcWidth;
        k = r % kNrcWidth;
    } else {
        layer = kNrcHidden;
        i = (p - kNrcOffOut) / kNrcWidth;
        k = (p - kNrcOffOut) % kNrcWidth;
    }
    float w = master[p];
    if (layer != kNrcHidden || i < 3) {
        float gr = nrcSumParts(grad, parts, p) * scale;
        float a = b1 * m1[p] + (1.0f - b1) * gr;
        float b = b2 * m2[p] + (1.0f - b2) * gr * gr;
        m1[p] = a;
        m2[p] = b;
        w -= lr * (a / c1) / (sqrtf(b / c2) + eps);
        master[p] = w;
    }
    const _Float16 hw = (_Float16)w;
    if (layer == 0) {
        wfwd[p] = hw;
    } else if (layer < kNrcHidden) {
        const int base = kNrcOffHidden + (layer - 1) * kNrcWidth * kNrcWidth;
        wfwd[base + i * kNrcWidth + nrcSlotOfNeuron(k)] = hw;
        wbwd[(layer - 1) * kNrcWidth * kNrcWidth + k * kNrcWidth + nrcSlotOfNeuron(i)] = hw;
    } else {
        wfwd[kNrcOffOut + i * kNrcWidth + nrcSlotOfNeuron(k)] = hw;
        wbwd[kNrcBwdOut + k * kNrcOutRows + i] = hw;
    }
}

}  // namespace mrt

using namespace mrt;

// ----------------------------------------------------------------- host ----
struct NrcState {
    DeviceArray<float> master, m1, m2, grad, gradSum, loss;  // grad: one partial per chunk of a batch
    DeviceArray<_Float16> wfwd, wbwd;
    DeviceArray<_Float16> tab;  // layer inputs and deltas of one slice of a batch
    DeviceArray<float4> scratchIn, scratchTarget, scratchOut;
    uint32_t step = 0;
    uint32_t inferGrid = 0;
};

namespace mrt {
static NrcState* nrcState(mrt_context* ctx) { return reinterpret_cast<NrcState*>(ctx->nrc); }

// Uniform initialisation with bound sqrt(6 / fan_in) (He et al. for ReLU
// layers), a tenth of it for the output layer; splitmix64 stream from seed.
static void nrcInitWeights(std::vector<float>& w, uint32_t seed) {
    uint64_t s = 0x9E3779B97F4A7C15ull * (seed + 1u);
    auto next = [&]() {
        s += 0x9E3779B97F4A7C15ull;
        uint64_t z = s;
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        z ^= z >> 31;
        return (float)((z >> 40) * (1.0 / 16777216.0)) * 2.0f - 1.0f;
    };
    w.assign(kNrcParams, 0.0f);
    const float b0 = sqrtf(6.0f / 173.0f), bh = sqrtf(6.0f / kNrcWidth);
    for (int i = 0; i < kNrcWidth; i++)
        for (int k = 0; k < 173; k++) w[kNrcOffW0 + i * kNrcIn + k] = b0 * next();
    for (int p = kNrcOffHidden; p < kNrcOffOut; p++) w[p] = bh * next();
    for (int i = 0; i < 3; i++)
        for (int k = 0; k < kNrcWidth; k++) w[kNrcOffOut + i * kNrcWidth + k] = 0.1f * bh * next();
}

static bool nrcUploadMaster(mrt_context* ctx, const float* w) {
    NrcState* st = nrcState(ctx);
    if (!st->master.upload(w, kNrcParams, ctx->stream)) return false;
    // rebuild the fp16 copies through one zero-step Adam pass (lr 0 keeps the weights)
    hipLaunchKernelGGL(kNrcAdam, dim3((kNrcParams + kNrcBlock - 1) / kNrcBlock), dim3(kNrcBlock), 0, ctx->stream,
                       st->master.ptr, st->m1.ptr, st->m2.ptr, st->grad.ptr, 0u, st->wfwd.ptr, st->wbwd.ptr, 0.0f, 0.0f,
                       0.9f, 0.99f, 1.0f, 1.0f, 1e-8f);
    return check(hipGetLastError(), "nrc weights");
}

bool nrcEnsure(mrt_context* ctx, uint32_t seed) {
    if (ctx->nrc) return true;
    NrcState* st = new NrcState();
    ctx->nrc = st;
    st->inferGrid = (uint32_t)std::max(1, ctx->cuCount) * 2u;
    bool ok = st->master.reserve(kNrcParams) && st->m1.reserve(kNrcParams) && st->m2.reserve(kNrcParams) &&
              st->loss.reserve(1) &&
              st->wfwd.reserve(kNrcParams) && st->wbwd.reserve(kNrcBwdParams) && st->gradSum.reserve(kNrcParams);
    if (!ok) return false;
    hipMemsetAsync(st->m1.ptr, 0, kNrcParams * sizeof(float), ctx->stream);
    hipMemsetAsync(st->m2.ptr, 0, kNrcParams * sizeof(float), ctx->stream);
    hipMemsetAsync(st->wbwd.ptr, 0, kNrcBwdParams * sizeof(_Float16), ctx->stream);
    std::vector<float> w;
    nrcInitWeights(w, seed);
    if (!nrcUploadMaster(ctx, w.data())) return false;
    if (!check(hipFuncSetAttribute(reinterpret_cast<const void*>(kNrcBackprop),
                                   hipFuncAttributeMaxDynamicSharedMemorySize, (int)kNrcTrainLds),
               "nrc training LDS"))
        return false;
    return check(hipStreamSynchronize(ctx->stream), "nrc init");
}

void nrcRelease(mrt_context* ctx) {
    NrcState* st = nrcState(ctx);
    if (!st) return;
    st->master.release();
    st->m1.release();
    st->m2.release();
    st->grad.release();
    st->gradSum.release();
    st->tab.release();
    st->loss.release();
    st->wfwd.release();
    st->wbwd.release();
    st->scratchIn.release();
    st->scratchTarget.release();
    st->scratchOut.release();
    delete st;
    ctx->nrc = nullptr;
}

// rad[path] += beta * cache(query) for *count queries (device counter)
// rad[path] += throughput * cache(query) for min(*count, capacity) queries of `stride` float4
bool nrcInferQueries(mrt_context* ctx, const float4* queries, uint32_t stride, const uint32_t* count,
                     uint32_t capacity, float4* rad, const Spec* spec) {
    NrcState* st = nrcState(ctx);
    hipLaunchKernelGGL(kNrcInfer, dim3(st->inferGrid), dim3(kNrcBlock), 0, ctx->stream, st->wfwd.ptr, queries, stride,
                       count, capacity, nullptr, rad, spec ? *spec : Spec{}, spec ? 1u : 0u);
    return check(hipGetLastError(), "nrc inference");
}

// The two training passes over a batch of `batch` slots (slot b takes record
// hash(batchOffset + b) % n when shuffle is set, else b % n, of the n =
// min(*count, capacity) records), in slices of kNrcSlice slots. Leaves the
// gradient as partials in st->grad and returns their number (0 on failure).
static uint32_t nrcGradientPasses(mrt_context* ctx, const float4* in, uint32_t stride, const float4* target,
                                  const uint32_t* count, uint32_t capacity, uint32_t batch, uint32_t batchOffset,
                                  uint32_t shuffle) {
    NrcState* st = nrcState(ctx);
    const uint32_t parts = (std::min(batch, kNrcSlice) + kNrcChunk - 1) / kNrcChunk;
    if (!st->tab.reserve((size_t)kNrcTabRows * parts * kNrcChunk) || !st->grad.reserve((size_t)parts * kNrcParams))
        return 0;
    for (uint64_t b = 0; b < batch; b += kNrcSlice) {
        const uint32_t len = (uint32_t)std::min<uint64_t>(kNrcSlice, batch - b);
        const uint32_t chunks = (len + kNrcChunk - 1) / kNrcChunk, ld = chunks * kNrcChunk;
        hipLaunchKernelGGL(kNrcBackprop, dim3(ld / (kNrcTrainBlock / 4)), dim3(kNrcTrainBlock), kNrcTrainLds, ctx->stream, st->wfwd.ptr,
                           st->wbwd.ptr,
                           in, stride, target, count, capacity, batch, (uint32_t)b, batchOffset, shuffle, st->tab.ptr,
                           ld, st->loss.ptr);
        // later slices add to the first slice's partials
        hipLaunchKernelGGL(kNrcWgrad, dim3((chunks * kNrcWgradTasks + 3) / 4), dim3(kNrcBlock), 0, ctx->stream,
                           st->tab.ptr, ld, chunks, st->grad.ptr, b > 0 ? 1u : 0u);
    }
    return check(hipGetLastError(), "nrc gradient passes") ? parts : 0u;
}

// `steps` Adam steps of `batch` records drawn from the min(*count, capacity)
// training records (count may be null: capacity records)
bool nrcTrainSteps(mrt_context* ctx, const float4* in, uint32_t stride, const float4* target, const uint32_t* count,
                   uint32_t capacity, uint32_t steps, uint32_t batch, float lr) {
    NrcState* st = nrcState(ctx);
    if (capacity == 0 || steps == 0 || batch == 0) return true;
    hipMemsetAsync(st->loss.ptr, 0, sizeof(float), ctx->stream);
    for (uint32_t s = 0; s < steps; s++) {
        const uint32_t parts = nrcGradientPasses(ctx, in, stride, target, count, capacity, batch, st->step * batch, 1u);
        if (parts == 0) return false;
        st->step++;
        const float b1 = 0.9f, b2 = 0.99f;
        const float c1 = 1.0f - powf(b1, (float)st->step), c2 = 1.0f - powf(b2, (float)st->step);
        hipLaunchKernelGGL(kNrcAdam, dim3((kNrcParams + kNrcBlock - 1) / kNrcBlock), dim3(kNrcBlock), 0, ctx->stream,
                           st->master.ptr, st->m1.ptr, st->m2.ptr, st->grad.ptr, parts, st->wfwd.ptr,
                           st->wbwd.ptr,
                           1.0f / (float)batch, lr, b1, b2, c1, c2, 1e-8f);
    }
    return check(hipGetLastError(), "nrc training");
}

float nrcReadLoss(mrt_context* ctx) {
    NrcState* st = nrcState(ctx);
    float v = 0.0f;
    if (!st || !check(hipMemcpyAsync(&v, st->loss.ptr, sizeof(float), hipMemcpyDeviceToHost, ctx->stream), "nrc loss"))
        return -1.0f;
    hipStreamSynchronize(ctx->stream);
    return v;
}
}  // namespace mrt

// ------------------------------------------------------------- C ABI ------
extern "C" int mrt_nrc_reset(mrt_context* ctx, uint32_t seed) {
    if (!ctx) return -1;
    if (!check(hipSetDevice(ctx->device), "hipSetDevice")) return -2;
    nrcRelease(ctx);
    return nrcEnsure(ctx, seed) ? 0 : -3;
}

extern "C" int mrt_nrc_parameter_count(void) { return kNrcParams; }

extern "C" int mrt_nrc_get_weights(mrt_context* ctx, float* out, size_t count) {
    if (!ctx || !out || count != (size_t)kNrcParams) return -1;
    if (!nrcEnsure(ctx, 0)) return -3;
    NrcState* st = nrcState(ctx);
    if (!check(hipMemcpyAsync(out, st->master.ptr, kNrcParams * sizeof(float), hipMemcpyDeviceToHost, ctx->stream),
               "nrc weights D2H"))
        return -3;
    return check(hipStreamSynchronize(ctx->stream), "nrc weights") ? 0 : -3;
}

extern "C" int mrt_nrc_set_weights(mrt_context* ctx, const float* w, size_t count) {
    if (!ctx || !w || count != (size_t)kNrcParams) return -1;
    if (!nrcEnsure(ctx, 0)) return -3;
    std::vector<float> v(w, w + kNrcParams);
    for (int i = 3; i < kNrcOutRows; i++)
        for (int k = 0; k < kNrcWidth; k++) v[kNrcOffOut + i * kNrcWidth + k] = 0.0f;
    if (!nrcUploadMaster(ctx, v.data())) return -3;
    return check(hipStreamSynchronize(ctx->stream), "nrc set weights") ? 0 : -3;
}

extern "C" int mrt_nrc_infer(mrt_context* ctx, const float* inputs, uint32_t n, float* outputs) {
    if (!ctx || (n && (!inputs || !outputs))) return -1;
    if (!nrcEnsure(ctx, 0)) return -3;
    if (n == 0) return 0;
    NrcState* st = nrcState(ctx);
    std::vector<float4> in(3 * (size_t)n);
    for (uint32_t i = 0; i < n; i++) {
        const float* v = inputs + 11 * (size_t)i;
        in[3 * i] = make_float4(v[0], v[1], v[2], v[3]);
        in[3 * i + 1] = make_float4(v[4], v[5], v[6], v[7]);
        in[3 * i + 2] = make_float4(v[8], v[9], v[10], 0.0f);
    }
    if (!st->scratchIn.upload(in.data(), in.size(), ctx->stream) || !st->scratchOut.reserve(n)) return -3;
    hipLaunchKernelGGL(kNrcInfer, dim3(st->inferGrid), dim3(kNrcBlock), 0, ctx->stream, st->wfwd.ptr,
                       st->scratchIn.ptr, 3u, nullptr, n, st->scratchOut.ptr, nullptr, Spec{}, 0u);
    if (!check(hipGetLastError(), "nrc inference")) return -3;
    std::vector<float4> out(n);
    if (!check(hipMemcpyAsync(out.data(), st->scratchOut.ptr, n * sizeof(float4), hipMemcpyDeviceToHost, ctx->stream),
               "nrc outputs D2H") ||
        !check(hipStreamSynchronize(ctx->stream), "nrc inference"))
        return -3;
    for (uint32_t i = 0; i < n; i++) {
        outputs[3 * i] = out[i].x;
        outputs[3 * i + 1] = out[i].y;
        outputs[3 * i + 2] = out[i].z;
    }
    return 0;
}

extern "C" int mrt_nrc_train(mrt_context* ctx, const float* inputs, const float* targets, uint32_t n, uint32_t steps,
                             uint32_t batch, float learningRate, float* lossOut) {
    if (!ctx || (n && (!inputs || !targets))) return -1;
    if (!nrcEnsure(ctx, 0)) return -3;
    NrcState* st = nrcState(ctx);
    std::vector<float4> in(3 * (size_t)n), tg(n);
    for (uint32_t i = 0; i < n; i++) {
        const float* v = inputs + 11 * (size_t)i;
        in[3 * i] = make_float4(v[0], v[1], v[2], v[3]);
        in[3 * i + 1] = make_float4(v[4], v[5], v[6], v[7]);
        in[3 * i + 2] = make_float4(v[8], v[9], v[10], 0.0f);
        tg[i] = make_float4(targets[3 * i], targets[3 * i + 1], targets[3 * i + 2], 0.0f);
    }
    if (!st->scratchIn.upload(in.data(), in.size(), ctx->stream) ||
        !st->scratchTarget.upload(tg.data(), tg.size(), ctx->stream))
        return -3;
    if (!nrcTrainSteps(ctx, st->scratchIn.ptr, 3u, st->scratchTarget.ptr, nullptr, n, steps, batch, learningRate))
        return -3;
    float loss = nrcReadLoss(ctx);
    if (lossOut) *lossOut = loss;
    return loss < 0.0f ? -3 : 0;
}

// Gradient of the summed relative L2 loss over the n given samples (each once),
// without an optimiser step: grad receives kNrcParams floats.
extern "C" int mrt_nrc_gradient(mrt_context* ctx, const float* inputs, const float* targets, uint32_t n, float* grad,
                                float* lossOut) {
    if (!ctx || !inputs || !targets || !grad || n == 0) return -1;
    if (!nrcEnsure(ctx, 0)) return -3;
    NrcState* st = nrcState(ctx);
    std::vector<float4> in(3 * (size_t)n), tg(n);
    for (uint32_t i = 0; i < n; i++) {
        const float* v = inputs + 11 * (size_t)i;
        in[3 * i] = make_float4(v[0], v[1], v[2], v[3]);
        in[3 * i + 1] = make_float4(v[4], v[5], v[6], v[7]);
        in[3 * i + 2] = make_float4(v[8], v[9], v[10], 0.0f);
        tg[i] = make_float4(targets[3 * i], targets[3 * i + 1], targets[3 * i + 2], 0.0f);
    }
    if (!st->scratchIn.upload(in.data(), in.size(), ctx->stream) ||
        !st->scratchTarget.upload(tg.data(), tg.size(), ctx->stream))
        return -3;
    hipMemsetAsync(st->loss.ptr, 0, sizeof(float), ctx->stream);
    const uint32_t parts = nrcGradientPasses(ctx, st->scratchIn.ptr, 3u, st->scratchTarget.ptr, nullptr, n, n, 0u, 0u);
    if (parts == 0) return -3;
    hipLaunchKernelGGL(kNrcSumParts, dim3((kNrcParams + kNrcBlock - 1) / kNrcBlock), dim3(kNrcBlock), 0, ctx->stream,
                       st->grad.ptr, parts, st->gradSum.ptr);
    if (!check(hipGetLastError(), "nrc gradient")) return -3;
    if (!check(hipMemcpyAsync(grad, st->gradSum.ptr, kNrcParams * sizeof(float), hipMemcpyDeviceToHost, ctx->stream),
               "nrc gradient D2H"))
        return -3;
    float loss = nrcReadLoss(ctx);
    if (lossOut) *lossOut = loss;
    return loss < 0.0f ? -3 : 0;
}
