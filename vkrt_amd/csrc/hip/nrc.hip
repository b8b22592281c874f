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
// LDS copy of the gradient during training, rows padded by 4 floats so that the
// four lane groups of a fragment (rows 4 apart) fall in different banks
constexpr int kNrcPadIn = kNrcIn + 4, kNrcPadW = kNrcWidth + 4;
constexpr int kNrcLdsHidden = kNrcWidth * kNrcPadIn;                          // 12544
constexpr int kNrcLdsOut = kNrcLdsHidden + 4 * kNrcWidth * kNrcPadW;          // 29952
constexpr int kNrcLdsGrad = kNrcLdsOut + kNrcOutRows * kNrcWidth;             // 30976

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

// Forward pass of 16 samples (one per lane%16). w: forward fp16 weights (LDS or global).
// Saves the hidden activations as B fragments when h != nullptr. Returns the output D fragment.
template <bool SAVE>
__device__ inline float4v nrcForward(const _Float16* w, const half8 x[6], half8 (*h)[2]) {
    const int l = threadIdx.x & 63, row = l & 15, g = l >> 4;
    float4v d[4];
#pragma unroll
    for (int ib = 0; ib < 4; ib++) {
        float4v acc = {0.0f, 0.0f, 0.0f, 0.0f};
        const _Float16* wr = w + kNrcOffW0 + (ib * 16 + row) * kNrcIn + 8 * g;
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
        const _Float16* wl = w + kNrcOffHidden + (layer - 1) * kNrcWidth * kNrcWidth;
#pragma unroll
        for (int ib = 0; ib < 4; ib++) {
            float4v acc = {0.0f, 0.0f, 0.0f, 0.0f};
            const _Float16* wr = wl + (ib * 16 + row) * kNrcWidth + 8 * g;
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
    const _Float16* wo = w + kNrcOffOut + row * kNrcWidth + 8 * g;
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
    __shared__ __align__(16) _Float16 w[kNrcParams];
    for (int k = threadIdx.x; k < kNrcParams / 8; k += kNrcBlock)
        reinterpret_cast<half8*>(w)[k] = reinterpret_cast<const half8*>(wfwd)[k];
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

// 16x16 table of one wave in LDS, [row][16 samples]
__device__ inline void nrcWaitLds() {
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
}

// One optimiser step's gradient: every wave takes blocks of 16 samples of the
// batch, runs the forward pass keeping the activations, back-propagates the
// relative L2 loss sum_c (y_c - t_c)^2 / (y_c^2 + 0.01) (y detached in the
// denominator) and adds the weight gradients into an LDS copy of the gradient,
// which each block writes out as its partial gradient (grad + block * params).
__global__ __launch_bounds__(kNrcBlock) void kNrcTrain(const _Float16* __restrict__ wfwd,
                                                      const _Float16* __restrict__ wbwd, const float4* __restrict__ in,
                                                      uint32_t stride, const float4* __restrict__ target,
                                                      const uint32_t* count, uint32_t nFixed, uint32_t batch,
                                                      uint32_t batchOffset, uint32_t shuffle, float* grad,
                                                      float* lossSum) {
    const uint32_t n = count ? min(*count, nFixed) : nFixed;
    extern __shared__ __align__(16) float nrcShared[];
    float* gl = nrcShared;                                                    // kNrcLdsGrad floats
    _Float16* tabBase = reinterpret_cast<_Float16*>(nrcShared + kNrcLdsGrad);  // per wave: 64x16 + 192x16 halfs
    for (int k = threadIdx.x; k < kNrcLdsGrad; k += kNrcBlock) gl[k] = 0.0f;
    __syncthreads();
    const int l = threadIdx.x & 63, row = l & 15, g = l >> 4;
    _Float16* td = tabBase + (threadIdx.x >> 6) * (kNrcWidth * 16 + kNrcIn * 16);  // deltas [64][16]
    _Float16* tx = td + kNrcWidth * 16;                                            // inputs [192][16]
    const uint32_t wave = blockIdx.x * (kNrcBlock / 64) + (threadIdx.x >> 6);
    const uint32_t waves = gridDim.x * (kNrcBlock / 64);
    float lossAcc = 0.0f;
    for (uint32_t base = wave * 16u; base < batch; base += waves * 16u) {
        const uint32_t bi = base + (uint32_t)row;
        const bool valid = bi < batch && n > 0;
        // a pseudo-random training record for this batch slot, or record bi itself
        const uint32_t idx = !valid ? 0u
                                    : (shuffle ? (uint32_t)(((uint64_t)(batchOffset + bi) * 2654435761ull) % n)
                                               : bi % n);
        float v[11];
        nrcLoadRaw(in, stride, idx, valid, v);
        half8 x[6];
        nrcEncode(v, g, x);
        half8 h[kNrcHidden][2];
        const int z = nrcOpaqueZero();
        const _Float16* wf = wfwd + z;
        const _Float16* wb = wbwd + z;
        float4v o = nrcForward<true>(wf, x, h);

        // loss gradient, D layout (rows 0..2 of lane group 0)
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

        // dW5 = dOut x h5^T over the 16 samples
#pragma unroll
        for (int t = 0; t < 4; t++) td[(4 * g + t) * 16 + row] = dOutH[t];
#pragma unroll
        for (int kb = 0; kb < 2; kb++)
#pragma unroll
            for (int t = 0; t < 8; t++) tx[nrcNeuronOfSlot(32 * kb + 8 * g + t) * 16 + row] = h[4][kb][t];
        nrcWaitLds();
        {
            half4 a = ldHalf4(td + row * 16 + 4 * g);
#pragma unroll
            for (int kc = 0; kc < 4; kc++) {
                half4 b = ldHalf4(tx + (kc * 16 + row) * 16 + 4 * g);
                float4v acc = __builtin_amdgcn_mfma_f32_16x16x16f16(a, b, float4v{0.0f, 0.0f, 0.0f, 0.0f}, 0, 0, 0);
                if (g == 0)
#pragma unroll
                    for (int t = 0; t < 3; t++) atomicAdd(&gl[kNrcLdsOut + t * kNrcWidth + kc * 16 + row], acc[t]);
            }
        }
        // delta of h5: (W5^T dOut) masked by h5 > 0, D layout in natural neuron blocks
        float4v dl[4];
#pragma unroll
        for (int mb = 0; mb < 4; mb++) {
            half4 a = ldHalf4(wb + kNrcBwdOut + (mb * 16 + row) * kNrcOutRows + 4 * g);
            float4v acc = __builtin_amdgcn_mfma_f32_16x16x16f16(a, dOutH, float4v{0.0f, 0.0f, 0.0f, 0.0f}, 0, 0, 0);
#pragma unroll
            for (int t = 0; t < 4; t++) acc[t] = (float)h[4][mb >> 1][(mb & 1) * 4 + t] > 0.0f ? acc[t] : 0.0f;
            dl[mb] = acc;
        }
        // hidden layers 4..1: dW_l = delta_{l+1} x h_l^T; delta_l = (W_l^T delta_{l+1}) * (h_l > 0)
#pragma unroll
        for (int layer = kNrcHidden - 1; layer >= 1; layer--) {
            nrcWaitLds();
#pragma unroll
            for (int mb = 0; mb < 4; mb++)
#pragma unroll
                for (int t = 0; t < 4; t++) td[(mb * 16 + 4 * g + t) * 16 + row] = (_Float16)dl[mb][t];
#pragma unroll
            for (int kb = 0; kb < 2; kb++)
#pragma unroll
                for (int t = 0; t < 8; t++)
                    tx[nrcNeuronOfSlot(32 * kb + 8 * g + t) * 16 + row] = h[layer - 1][kb][t];
            nrcWaitLds();
            float* gw = gl + kNrcLdsHidden + (layer - 1) * kNrcWidth * kNrcPadW;
#pragma unroll
            for (int ib = 0; ib < 4; ib++) {
                half4 a = ldHalf4(td + (ib * 16 + row) * 16 + 4 * g);
#pragma unroll
                for (int kc = 0; kc < 4; kc++) {
                    half4 b = ldHalf4(tx + (kc * 16 + row) * 16 + 4 * g);
                    float4v acc = __builtin_amdgcn_mfma_f32_16x16x16f16(a, b, float4v{0.0f, 0.0f, 0.0f, 0.0f}, 0, 0, 0);
#pragma unroll
                    for (int t = 0; t < 4; t++) atomicAdd(&gw[(ib * 16 + 4 * g + t) * kNrcPadW + kc * 16 + row], acc[t]);
                }
            }
            const half8 bd[2] = {nrcPack(dl[0], dl[1]), nrcPack(dl[2], dl[3])};
            const _Float16* wbl = wb + (layer - 1) * kNrcWidth * kNrcWidth;
#pragma unroll
            for (int mb = 0; mb < 4; mb++) {
                const _Float16* wr = wbl + (mb * 16 + row) * kNrcWidth + 8 * g;
                float4v acc = {0.0f, 0.0f, 0.0f, 0.0f};
                acc = __builtin_amdgcn_mfma_f32_16x16x32_f16(ldHalf8(wr), bd[0], acc, 0, 0, 0);
                acc = __builtin_amdgcn_mfma_f32_16x16x32_f16(ldHalf8(wr + 32), bd[1], acc, 0, 0, 0);
#pragma unroll
                for (int t = 0; t < 4; t++)
                    acc[t] = (float)h[layer - 1][mb >> 1][(mb & 1) * 4 + t] > 0.0f ? acc[t] : 0.0f;
                dl[mb] = acc;
            }
        }
        // dW0 = delta_1 x x^T
        nrcWaitLds();
#pragma unroll
        for (int mb = 0; mb < 4; mb++)
#pragma unroll
            for (int t = 0; t < 4; t++) td[(mb * 16 + 4 * g + t) * 16 + row] = (_Float16)dl[mb][t];
#pragma unroll
        for (int kb = 0; kb < 6; kb++)
#pragma unroll
            for (int t = 0; t < 8; t++) tx[(32 * kb + 8 * g + t) * 16 + row] = x[kb][t];
        nrcWaitLds();
#pragma unroll
        for (int ib = 0; ib < 4; ib++) {
            half4 a = ldHalf4(td + (ib * 16 + row) * 16 + 4 * g);
            for (int kc = 0; kc < kNrcIn / 16; kc++) {
                half4 b = ldHalf4(tx + (kc * 16 + row) * 16 + 4 * g);
                float4v acc = __builtin_amdgcn_mfma_f32_16x16x16f16(a, b, float4v{0.0f, 0.0f, 0.0f, 0.0f}, 0, 0, 0);
#pragma unroll
                for (int t = 0; t < 4; t++) atomicAdd(&gl[(ib * 16 + 4 * g + t) * kNrcPadIn + kc * 16 + row], acc[t]);
            }
        }
        nrcWaitLds();
    }
    __syncthreads();
    // this block's partial gradient, summed over blocks by the optimiser step
    float* part = grad + (size_t)blockIdx.x * kNrcParams;
    for (int k = threadIdx.x; k < kNrcParams; k += kNrcBlock) {
        int src;
        if (k < kNrcOffHidden) {
            src = (k / kNrcIn) * kNrcPadIn + k % kNrcIn;
        } else if (k < kNrcOffOut) {
            const int r = k - kNrcOffHidden, layer = r / (kNrcWidth * kNrcWidth), e = r % (kNrcWidth * kNrcWidth);
            src = kNrcLdsHidden + layer * kNrcWidth * kNrcPadW + (e / kNrcWidth) * kNrcPadW + e % kNrcWidth;
        } else {
            src = kNrcLdsOut + (k - kNrcOffOut);
        }
        part[k] = gl[src];
    }
    // one loss sum per wave
    for (int off = 32; off > 0; off >>= 1) lossAcc += __shfl_down(lossAcc, off);
    if (l == 0 && lossAcc != 0.0f) atomicAdd(lossSum, lossAcc);
}

// Sum over the training blocks' partial gradients of parameter p, with 8
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

// Sum of the training blocks' partial gradients.
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
    DeviceArray<float> master, m1, m2, grad, gradSum, loss;  // grad: one partial per training block
    DeviceArray<_Float16> wfwd, wbwd;
    DeviceArray<float4> scratchIn, scratchTarget, scratchOut;
    uint32_t step = 0;
    uint32_t inferGrid = 0, trainGrid = 0;
};

namespace mrt {
size_t nrcTrainLds() { return (size_t)kNrcLdsGrad * 4 + (size_t)(kNrcBlock / 64) * (kNrcWidth * 16 + kNrcIn * 16) * 2; }

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
    // one training block per CU (the LDS gradient copy fills most of a CU's LDS)
    st->inferGrid = (uint32_t)std::max(1, ctx->cuCount) * 2u;
    st->trainGrid = (uint32_t)std::max(1, ctx->cuCount);
    bool ok = st->master.reserve(kNrcParams) && st->m1.reserve(kNrcParams) && st->m2.reserve(kNrcParams) &&
              st->grad.reserve((size_t)st->trainGrid * kNrcParams) && st->loss.reserve(1) &&
              st->wfwd.reserve(kNrcParams) && st->wbwd.reserve(kNrcBwdParams) && st->gradSum.reserve(kNrcParams);
    if (!ok) return false;
    hipMemsetAsync(st->m1.ptr, 0, kNrcParams * sizeof(float), ctx->stream);
    hipMemsetAsync(st->m2.ptr, 0, kNrcParams * sizeof(float), ctx->stream);
    hipMemsetAsync(st->wbwd.ptr, 0, kNrcBwdParams * sizeof(_Float16), ctx->stream);
    std::vector<float> w;
    nrcInitWeights(w, seed);
    if (!nrcUploadMaster(ctx, w.data())) return false;
    if (!check(hipFuncSetAttribute(reinterpret_cast<const void*>(kNrcTrain),
                                   hipFuncAttributeMaxDynamicSharedMemorySize, (int)nrcTrainLds()),
               "nrc train LDS"))
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

// `steps` Adam steps of `batch` records drawn from the min(*count, capacity)
// training records (count may be null: capacity records)
bool nrcTrainSteps(mrt_context* ctx, const float4* in, uint32_t stride, const float4* target, const uint32_t* count,
                   uint32_t capacity, uint32_t steps, uint32_t batch, float lr) {
    NrcState* st = nrcState(ctx);
    if (capacity == 0 || steps == 0 || batch == 0) return true;
    hipMemsetAsync(st->loss.ptr, 0, sizeof(float), ctx->stream);
    for (uint32_t s = 0; s < steps; s++) {
        hipLaunchKernelGGL(kNrcTrain, dim3(st->trainGrid), dim3(kNrcBlock), nrcTrainLds(), ctx->stream, st->wfwd.ptr,
                           st->wbwd.ptr, in, stride, target, count, capacity, batch, st->step * batch, 1u,
                           st->grad.ptr, st->loss.ptr);
        st->step++;
        const float b1 = 0.9f, b2 = 0.99f;
        const float c1 = 1.0f - powf(b1, (float)st->step), c2 = 1.0f - powf(b2, (float)st->step);
        hipLaunchKernelGGL(kNrcAdam, dim3((kNrcParams + kNrcBlock - 1) / kNrcBlock), dim3(kNrcBlock), 0, ctx->stream,
                           st->master.ptr, st->m1.ptr, st->m2.ptr, st->grad.ptr, st->trainGrid, st->wfwd.ptr,
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
    hipLaunchKernelGGL(kNrcTrain, dim3(st->trainGrid), dim3(kNrcBlock), nrcTrainLds(), ctx->stream, st->wfwd.ptr,
                       st->wbwd.ptr, st->scratchIn.ptr, 3u, st->scratchTarget.ptr, nullptr, n, n, 0u, 0u,
                       st->grad.ptr, st->loss.ptr);
    hipLaunchKernelGGL(kNrcSumParts, dim3((kNrcParams + kNrcBlock - 1) / kNrcBlock), dim3(kNrcBlock), 0, ctx->stream,
                       st->grad.ptr, st->trainGrid, st->gradSum.ptr);
    if (!check(hipGetLastError(), "nrc gradient")) return -3;
    if (!check(hipMemcpyAsync(grad, st->gradSum.ptr, kNrcParams * sizeof(float), hipMemcpyDeviceToHost, ctx->stream),
               "nrc gradient D2H"))
        return -3;
    float loss = nrcReadLoss(ctx);
    if (lossOut) *lossOut = loss;
    return loss < 0.0f ? -3 : 0;
}
