// Wavefront path tracer for gfx950 -- the MI355X replacement for the
// reference's raygen megakernels (src/shaders/entry/path/raygen_*.slang) and
// their per-pixel loop (src/shaders/integrator/path/{rgb,spectral_single,
// spectral_hero}/integrator.slang).
//
// One frame = samplesPerPixel paths per viewport pixel, processed in chunks of
// at most kMaxChunk paths. Path p of the frame is (sample s, pixel k) with
// p = s * pixels + k, so a chunk of whole samples touches every pixel once per
// sample and consecutive lanes shade neighbouring pixels. Per chunk:
//   generate -> [ extend (closest hit) -> shade -> shadow (any hit) ] x depth
//   -> finalize (per-path XYZ / RGB) -> reduce (per-pixel sum in sample order)
// and once per frame: accumulate (running mean, display encode).
// Queues live in HBM and are compacted with one atomic per wavefront.
// Shadow rays carry two contributions (visible / not visible) so the
// radiance update is the reference's `radiance += throughput * contribution`
// evaluated with the same operands once visibility is known.
#include <algorithm>
#include <chrono>
#include <cstring>

#include <rocprim/device/device_scan.hpp>

#include "dev_shading.h"
#include "dev_traverse.h"
#include "mrt_internal.hpp"

namespace mrt {

enum : int { kModeRgb = 0, kModeSingle = 1, kModeHero = 2 };

constexpr uint32_t kFPrevNee = 1u;
constexpr uint32_t kFHeroActive = 2u;
constexpr uint32_t kFRefractive = 4u;
constexpr uint32_t kFAbsorption = 8u;
constexpr uint32_t kFPendingCollapse = 16u;
constexpr uint32_t kFPrevEnvNee = 32u;  // the previous vertex sampled the environment for NEE

constexpr uint32_t kVisOccluded = 0u, kVisVisible = 1u, kVisUnsupported = 2u;

// Paths in flight per chunk. Large enough that the queues of late bounces
// (after most paths have ended) still fill the chip: 2^27 paths keep about
// 40 GB of path state and queues resident, well inside 288 GB of HBM (a
// 1080p frame of 64 samples is one chunk).
constexpr uint64_t kMaxChunk = 1ull << 27;
constexpr int kBlock = 256;

// counters[] layout
constexpr int kCntQueue0 = 0, kCntQueue1 = 1, kCntShadow = 2, kCntShadowEnv = 3, kCntNrcQuery = 4, kCntDropped = 5,
              kCntNrcRecord = 6, kCntNrcTrain = 7,
              kCntStats = 8;  // 64-bit stats at counters+8: [0] ext rays, [1] shadow rays, [2] nodes, [3] triangles,
                              // [4] rays of the fused tail, [5..7] [0], [2], [3] as they stood when the tail began
// queue length of every depth (the first kDepthHist depths) and the fused tail's claim counter
constexpr int kCntDepthHist = 24, kDepthHist = 32, kCntTailFetch = kCntDepthHist + kDepthHist;
// Work claims of the persistent traversal kernels. The batch space of a queue
// is split into kFetchParts interleaved partitions, each with its own counter
// on its own 256-byte line: device-scope atomics from the 8 XCDs then spread
// over 8 addresses instead of serialising on one. A wave claims kFetchBatch
// rays per atomic.
constexpr int kFetchParts = 8;
constexpr uint32_t kFetchStride = 64;  // uint32 slots between partition counters
constexpr uint32_t kFetchBatch = 64;
constexpr int kCntFetchExt = 64, kCntFetchShadow = kCntFetchExt + kFetchParts * kFetchStride,
              kCntFetchShadowEnv = kCntFetchShadow + kFetchParts * kFetchStride;
constexpr uint32_t kCounterWords = kCntFetchShadowEnv + kFetchParts * kFetchStride;

struct FrameParams {
    SceneData sd;
    uint64_t pathBegin;
    uint32_t count;
    uint32_t numPixels;
    uint32_t vpW;
    uint32_t accumWidth;
    uint32_t rngPolicy;
    uint64_t rngSeed;
    uint32_t neeEnabled, bsdfOnly, neeOnly;
    uint32_t depth;
    uint32_t traceShadowAlways;
    uint32_t envNee;  // environment importance sampling for NEE
    Spec spec;
    // neural radiance cache (RGB, and hero spectral with the cache in CIE XYZ): paths
    // [0, nrcRenderPaths) of the chunk render the image and query the cache after
    // nrcPrefix bounces; the rest are training paths
    uint32_t nrc, nrcPrefix, nrcRenderPaths;
    float nrcBoxMin[3], nrcBoxInv[3];
    // kGenerate wrote only the primary rays; kShade builds each path's state at depth 0
    uint32_t leanStart;
    // lean start without the feature pass: kGenerate stores only the primary rays'
    // directions; their common origin (the camera's) is recomputed where read
    uint32_t compactPrimary;
    const float4* accum;  // accumulation image (sample counts for the random streams)
};

// Neural radiance cache records. A query is 4 float4: the 11 network inputs in
// three (the third's w holds the path index) and the throughput. A training
// record is 6 float4: the inputs (w = path index), the throughput arriving at
// the vertex, the path radiance before the vertex and the emission it adds.
struct NrcView {
    float4* query;
    float4* record;
    float4* trainIn;
    float4* trainTarget;
    uint32_t queryCapacity, recordCapacity;
};
constexpr int kNrcQueryWords = 5, kNrcRecordWords = 7;

// What one chunk does with the neural cache after its depth loop.
struct NrcRun {
    NrcView view;
    bool train;
    uint32_t steps, batch;
    float learningRate;
};

struct PathView {
    uint4* rng;
    uint32_t* flags;
    float4 *thr, *rad, *radXYZ, *tpp, *ptpp, *pbtp, *lam, *misc, *sigma;
};

struct QueueView {
    float4* o;
    float4* d;
    uint32_t* count;
};

struct ShadowView {
    float4 *o, *d, *a, *b;
    uint32_t* count;
};

DEV Rng loadRng(uint4 v, uint32_t policy) {
    Rng r;
    r.policy = policy;
    r.state = ((uint64_t)v.y << 32) | v.x;
    r.inc = ((uint64_t)v.w << 32) | v.z;
    return r;
}
DEV uint4 storeRng(const Rng& r) {
    return make_uint4((uint32_t)r.state, (uint32_t)(r.state >> 32), (uint32_t)r.inc, (uint32_t)(r.inc >> 32));
}

// Appends this lane's item if pred; one atomic per wavefront. Every active lane
// that may append must call it (lanes that have left the calling branch do
// not take part). Returns the slot, or 0xFFFFFFFF.
DEV uint32_t waveAppend(bool pred, uint32_t* counter) {
    unsigned long long mask = __ballot(pred);
    if (mask == 0ull) return 0xFFFFFFFFu;
    uint32_t lane = __lane_id();
    int leader = __ffsll(mask) - 1;
    uint32_t base = 0;
    if ((int)lane == leader) base = atomicAdd(counter, (uint32_t)__popcll(mask));
    base = (uint32_t)__shfl((int)base, leader);
    return pred ? base + (uint32_t)__popcll(mask & ((1ull << lane) - 1ull)) : 0xFFFFFFFFu;
}

// First ray of the next batch of this wave's partition, or >= n once the
// partition is exhausted. Wave-uniform; every lane must call it.
DEV uint32_t claimBatch(uint32_t* fetchBase, uint32_t part, uint32_t n) {
    uint32_t k = 0;
    if ((threadIdx.x & 63u) == 0u) k = atomicAdd(&fetchBase[part * kFetchStride], 1u);
    k = (uint32_t)__shfl((int)k, 0);
    uint64_t first = ((uint64_t)k * kFetchParts + part) * kFetchBatch;
    return first >= n ? n : (uint32_t)first;
}

// ------------------------------------------------------------ lights --
// src/shaders/sampling/discrete.slang:8-18
DEV float powerHeuristic(float a, float b) {
    float a2 = a * a;
    return a2 / (a2 + b * b);
}
DEV uint32_t sampleAlias(float u, uint32_t count, uint32_t offset, const float* q, const uint32_t* idx) {
    float scaled = u * (float)count;
    uint32_t i = min((uint32_t)scaled, count - 1u);
    float rem = scaled - (float)i;
    return rem < q[offset + i] ? i : idx[offset + i];
}

struct LightSurfaceSample {
    F3 wi;
    F3 emission;
    float pdfSolidAngle, shadowDistance;
    uint32_t light;  // emissive mesh index
    bool valid;
};

// src/shaders/light/direct/light_sampling.slang:9-63
DEV LightSurfaceSample sampleLightSurface(const SceneView& sv, const SceneData& sd, F3 hitPoint, Rng& rng) {
    LightSurfaceSample s;
    s.valid = false;
    uint32_t meshCount = sd.emissiveMeshCount;
    if (meshCount == 0u) return s;
    uint32_t mi = sampleAlias(rnd(rng), meshCount, 0u, sv.meshAliasQ, sv.meshAliasIdx);
    const EmissiveMesh em = sv.emissiveMeshes[mi];
    s.light = mi;
    if (em.triCount == 0u) return s;
    uint32_t lt = sampleAlias(rnd(rng), em.triCount, em.triOffset, sv.triAliasQ, sv.triAliasIdx);
    const EmissiveTriangle tri = sv.emissiveTriangles[em.triOffset + lt];
    float u1 = rnd(rng);
    float u2 = rnd(rng);
    float su1 = sqrtf(u1);
    float b1 = u2 * su1;
    float b2 = (1.0f - u2) * su1;
    F3 v0 = F3{tri.v0Area[0], tri.v0Area[1], tri.v0Area[2]};
    F3 e1 = F3{tri.e1Pad[0], tri.e1Pad[1], tri.e1Pad[2]};
    F3 e2 = F3{tri.e2Pad[0], tri.e2Pad[1], tri.e2Pad[2]};
    F3 pos = v0 + e1 * b1 + e2 * b2;
    F3 nrm = safeNormalize(cross(e1, e2));
    s.emission = F3{em.emission[0], em.emission[1], em.emission[2]};
    float pdf = em.pmfMesh * em.invTotalArea;
    if (!(pdf > 0.0f)) return s;
    F3 toL = pos - hitPoint;
    float d2 = dot(toL, toL);
    if (d2 <= 0.0f) return s;
    float invD = rsq(d2);
    float dist = d2 * invD;
    s.wi = toL * invD;
    float cosL = fabsf(dot(s.wi, nrm));
    if (cosL <= 0.0f) return s;
    s.pdfSolidAngle = pdf * d2 / cosL;
    if (s.pdfSolidAngle <= 0.0f) return s;
    s.shadowDistance = dist - kShadowDistanceOffset;
    if (s.shadowDistance <= 0.0f) return s;
    s.valid = true;
    return s;
}

// ((x % W) + W) % W for x in [-W, 2W): the texel columns of the bilinear
// footprint, x0 = floor(u W - 1/2) in [-1, W - 1] and x1 = x0 + 1 in [0, W]
// for u in [0, 1], need one wrap at most (no integer division)
DEV int wrapOnce(int x, int W) { return x < 0 ? x + W : (x >= W ? x - W : x); }

// src/shaders/light/environment.slang:9-27 with the reference's environment
// sampler (linear filter, repeat U, clamp-to-edge V, level 0)
DEV F3 environmentRadiance(const SceneView& sv, const SceneData& sd, F3 dirIn) {
    if (sd.environmentTextureIndex == VKRT_INVALID_INDEX || !sv.env)
        return F3{sd.environmentLight[0], sd.environmentLight[1], sd.environmentLight[2]};
    F3 d = normalize(dirIn);
    float phi = atan2f(d.y, d.x) + sd.environmentRotation * (kPi / 180.0f);
    float theta = acosf(clampf(d.z, -1.0f, 1.0f));
    float u = fracf(phi * (0.5f * kInvPi) + 0.5f);
    float v = theta * kInvPi;
    int W = (int)sv.envWidth, H = (int)sv.envHeight;
    float fx = u * (float)W - 0.5f, fy = v * (float)H - 0.5f;
    float x0f = floorf(fx), y0f = floorf(fy);
    float tx = fx - x0f, ty = fy - y0f;
    int x0 = (int)x0f, y0 = (int)y0f;
    int x1 = x0 + 1, y1 = y0 + 1;
    x0 = wrapOnce(x0, W);
    x1 = wrapOnce(x1, W);
    y0 = min(max(y0, 0), H - 1);
    y1 = min(max(y1, 0), H - 1);
    F3 a00 = xyz(sv.env[(size_t)y0 * W + x0]), a10 = xyz(sv.env[(size_t)y0 * W + x1]);
    F3 a01 = xyz(sv.env[(size_t)y1 * W + x0]), a11 = xyz(sv.env[(size_t)y1 * W + x1]);
    F3 a = lerp(a00, a10, tx), b = lerp(a01, a11, tx);
    return lerp(a, b, ty) * sd.environmentLight[3];
}

// Environment importance sampling (an addition; the reference gathers the
// environment only on BSDF misses). Texel k is chosen with probability
// envPdf[k] (luminance x sin(theta) plus a floor, host table) through its alias
// table, the direction uniformly inside the texel in (u, v); the solid-angle pdf
// is envPdf[k] * W * H / (2 pi^2 sin(theta)). The (u, v) mapping is the inverse
// of environmentRadiance's.
struct EnvSample {
    F3 wi;
    float pdf;
    bool valid;
};
DEV EnvSample sampleEnvironmentDirection(const SceneView& sv, const SceneData& sd, Rng& rng) {
    EnvSample e;
    e.valid = false;
    const uint32_t W = sv.envWidth, H = sv.envHeight, n = W * H;
    float u1 = rnd(rng);
    float u2 = rnd(rng);
    float u3 = rnd(rng);
    float u4 = rnd(rng);
    uint32_t k = min((uint32_t)(u1 * (float)n), n - 1u);
    if (!(u2 < sv.envAliasQ[k])) k = sv.envAliasIdx[k];
    uint32_t x = k % W, y = k / W;
    float u = ((float)x + u3) / (float)W;
    float v = ((float)y + u4) / (float)H;
    float phi = (u - 0.5f) * (2.0f * kPi) - sd.environmentRotation * (kPi / 180.0f);
    float theta = v * kPi;
    float st = sinf(theta);
    if (!(st > 0.0f)) return e;
    e.wi = F3{st * cosf(phi), st * sinf(phi), cosf(theta)};
    e.pdf = sv.envPdf[k] * (float)n / (2.0f * kPi * kPi * st);
    e.valid = e.pdf > 0.0f;
    return e;
}
DEV float environmentPdf(const SceneView& sv, const SceneData& sd, F3 dirIn) {
    F3 d = normalize(dirIn);
    float phi = atan2f(d.y, d.x) + sd.environmentRotation * (kPi / 180.0f);
    float z = clampf(d.z, -1.0f, 1.0f);
    float u = fracf(phi * (0.5f * kInvPi) + 0.5f);
    float v = acosf(z) * kInvPi;
    const uint32_t W = sv.envWidth, H = sv.envHeight;
    uint32_t x = min((uint32_t)(u * (float)W), W - 1u), y = min((uint32_t)(v * (float)H), H - 1u);
    float st = sqrtf(fmaxf(0.0f, 1.0f - z * z));
    if (!(st > 0.0f)) return 0.0f;
    return sv.envPdf[(size_t)y * W + x] * (float)(W * H) / (2.0f * kPi * kPi * st);
}

// src/shaders/light/direct/mis_weights.slang:6-42
DEV float areaToSolidAngle(float pdfArea, F3 gn, F3 d, float t) {
    if (pdfArea <= 0.0f) return 0.0f;
    float d2 = t * t;
    if (d2 <= 0.0f) return 0.0f;
    float cosL = fabsf(dot(d, gn));
    if (cosL <= 0.0f) return 0.0f;
    return pdfArea * d2 / cosL;
}
DEV float spectralMis(F4 a, F4 b) {
    float den = sum(a) + sum(b);
    return den > 0.0f ? a.x / den : 0.0f;
}

DEV F3 transmittance3(F4 sigma, bool absorbing, float d) {
    if (!absorbing) return f3(1.0f);
    return F3{expf(-sigma.x * d), expf(-sigma.y * d), expf(-sigma.z * d)};
}
DEV F4 transmittance4(F4 sigma, bool absorbing, float d) {
    if (!absorbing) return f4(1.0f);
    return F4{expf(-sigma.x * d), expf(-sigma.y * d), expf(-sigma.z * d), expf(-sigma.w * d)};
}

// src/shaders/bsdf/base/medium.slang:75-113
DEV void updateMedium(const Spec& sp, const BMat& m, uint32_t ff, uint32_t isT, float l, uint32_t spectral,
                      uint32_t& flags, F4& sigma) {
    if (!isT) return;
    bool entering = ff != 0u;
    bool refr = entering && isRefractive(m);
    bool absb = entering && hasAbsorption(m);
    flags = (flags & ~(kFRefractive | kFAbsorption)) | (refr ? kFRefractive : 0u) | (absb ? kFAbsorption : 0u);
    F3 s3v = f3(0.0f);
    if (absb) {
        if (!spectral) {
            F3 t = maxs(sat(m.attenuationColor), 1e-6f);
            s3v = F3{-logf(t.x), -logf(t.y), -logf(t.z)} * m.absorptionCoefficient;
        } else {
            float t = fmaxf(spec1(sp, sat(m.attenuationColor), l), 1e-6f);
            s3v = f3(-logf(t) * m.absorptionCoefficient);
        }
    }
    sigma = spectral ? F4{s3v.x, s3v.x, s3v.x, s3v.x} : F4{s3v.x, s3v.y, s3v.z, s3v.x};
}
DEV void updateMedium4(const Spec& sp, const BMat& m, uint32_t ff, uint32_t isT, F4 l, uint32_t& flags, F4& sigma) {
    if (!isT) return;
    bool entering = ff != 0u;
    bool refr = entering && isRefractive(m);
    bool absb = entering && hasAbsorption(m);
    flags = (flags & ~(kFRefractive | kFAbsorption)) | (refr ? kFRefractive : 0u) | (absb ? kFAbsorption : 0u);
    if (!absb) {
        sigma = f4(0.0f);
        return;
    }
    F4 t = spec4(sp, sat(m.attenuationColor), l);
    t = F4{fmaxf(t.x, 1e-6f), fmaxf(t.y, 1e-6f), fmaxf(t.z, 1e-6f), fmaxf(t.w, 1e-6f)};
    sigma = F4{-logf(t.x), -logf(t.y), -logf(t.z), -logf(t.w)} * m.absorptionCoefficient;
}

// ------------------------------------------------------------ kernels --
// Pixel, random stream and wavelengths of chunk path i, as the primary ray was
// generated: image paths are sample s of pixel k with p = s * pixels + k; cache
// training paths go through a pseudo-random pixel on a stream of their own.
struct PathStart {
    Rng rng;    // after the two jitter draws and the wavelength draw
    F4 lam;
    float jx, jy;
    int px, py;
};
template <int MODE>
DEV PathStart pathStart(const FrameParams& fp, uint32_t i) {
    PathStart ps;
    uint64_t p = fp.pathBegin + i;
    uint32_t s = (uint32_t)(p / fp.numPixels);
    uint32_t k = (uint32_t)(p % fp.numPixels);
    const bool training = fp.nrc && i >= fp.nrcRenderPaths;
    if (training) {
        uint32_t h = (i - fp.nrcRenderPaths) * 0x9E3779B9u ^ (fp.sd.frameNumber * 0x85EBCA6Bu + 0x165667B1u);
        h ^= h >> 16;
        h *= 0x7FEB352Du;
        h ^= h >> 15;
        h *= 0x846CA68Bu;
        h ^= h >> 16;
        k = h % fp.numPixels;
    }
    ps.px = (int)(fp.sd.viewportRect[0] + k % fp.vpW);
    ps.py = (int)(fp.sd.viewportRect[1] + k / fp.vpW);
    float4 prev = fp.accum[(size_t)ps.py * fp.accumWidth + ps.px];
    uint32_t sampleAbs = training ? 0x80000000u + (i - fp.nrcRenderPaths) : (uint32_t)(prev.w + 0.5f) + s;
    ps.rng = rngInit(fp.rngPolicy, fp.rngSeed, ps.px, ps.py, fp.sd.frameNumber, sampleAbs);
    ps.jx = rnd(ps.rng);
    ps.jy = rnd(ps.rng);
    ps.lam = f4(0.0f);
    if (MODE == kModeSingle) {
        float u = wavelengthUnit(ps.rng, sampleAbs);
        ps.lam = F4{360.0f + sat(u) * 470.0f, 470.0f, 0.0f, 0.0f};
    } else if (MODE == kModeHero) {
        float u = wavelengthUnit(ps.rng, sampleAbs);
        ps.lam = F4{360.0f + fracf(u + 0.0f) * 470.0f, 360.0f + fracf(u + 0.25f) * 470.0f,
                    360.0f + fracf(u + 0.5f) * 470.0f, 360.0f + fracf(u + 0.75f) * 470.0f};
    }
    return ps;
}

// Primary rays of the chunk. With FULL the path state is initialised too (the
// adjoint's replay reads it); otherwise kShade rebuilds it at depth 0 from the
// same streams, which saves writing and re-reading about 130 bytes per path.
template <int MODE, bool FULL>
__global__ __launch_bounds__(kBlock) void kGenerate(FrameParams fp, PathView pv, QueueView q) {
    uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= fp.count) return;
    PathStart ps = pathStart<MODE>(fp, i);
    F3 o, d;
    primaryRay(fp.sd, ps.px, ps.py, ps.jx - 0.5f, ps.jy - 0.5f, o, d);
    if (FULL) {
        uint32_t flags = 0u;
        if (MODE != kModeRgb) pv.lam[i] = st4(ps.lam);
        if (MODE == kModeHero) {
            pv.radXYZ[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            pv.tpp[i] = make_float4(1.0f, 1.0f, 1.0f, 1.0f);
            // ptpp / pbtp are read only after a BSDF sample has written them
            flags |= kFHeroActive;
        }
        pv.rng[i] = storeRng(ps.rng);
        pv.flags[i] = flags;
        pv.thr[i] = make_float4(1.0f, 1.0f, 1.0f, 1.0f);
        pv.rad[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        // hero paths write every field of misc when they collapse to one wavelength,
        // before anything reads it
        if (MODE != kModeHero) pv.misc[i] = make_float4(0.0f, 1.0f, 0.0f, 0.0f);
        // sigma is read only inside an absorbing medium (kFAbsorption), after a transmission wrote it
    }
    if (FULL || !fp.compactPrimary) q.o[i] = make_float4(o.x, o.y, o.z, __uint_as_float(i));
    q.d[i] = make_float4(d.x, d.y, d.z, kRayTMax);
}

// The origin every primary ray has (primaryRay's org), from the same arithmetic.
DEV F3 cameraOrigin(const SceneData& sd) {
    F4 org = mulM4(sd.viewInverse, F4{0.0f, 0.0f, 0.0f, 1.0f});
    return F3{org.x, org.y, org.z};
}

__global__ void kPrepDepth(uint32_t* counters, int cur, int nxt, uint32_t depth) {
    const uint32_t t = threadIdx.x;
    if (t == 0) {
        unsigned long long* stats = reinterpret_cast<unsigned long long*>(counters + kCntStats);
        stats[0] += counters[cur];
        if (depth < (uint32_t)kDepthHist) counters[kCntDepthHist + depth] = counters[cur];
        counters[nxt] = 0u;
        counters[kCntShadow] = 0u;
        counters[kCntShadowEnv] = 0u;
        counters[kCntTailFetch] = 0u;
    }
    if (t < 3u * kFetchParts) counters[kCntFetchExt + t * kFetchStride] = 0u;
}
__global__ void kPrepShadow(uint32_t* counters) {
    unsigned long long* stats = reinterpret_cast<unsigned long long*>(counters + kCntStats);
    stats[1] += counters[kCntShadow] + counters[kCntShadowEnv];
}
__global__ void kSetCount(uint32_t* counters, int slot, uint32_t value) { counters[slot] = value; }
// The ray and traversal totals as they stand before the depth the fused tail takes over.
__global__ void kTailSnapshot(uint32_t* counters) {
    unsigned long long* stats = reinterpret_cast<unsigned long long*>(counters + kCntStats);
    stats[5] = stats[0];
    stats[6] = stats[2];
    stats[7] = stats[3];
}

// Closest-hit traversal, persistent with lane refill. Every loop iteration
// advances each busy lane by one node or leaf visit; once kRefillLanes lanes of
// a wavefront are idle, the wavefront claims that many new rays with one atomic,
// so lanes do not wait for the slowest ray of a batch. The queue is split into
// kFetchParts contiguous partitions, each with its own claim counter (its own
// 256-byte line); a wavefront starts in the partition of its block and moves on
// to the next partition once one is exhausted.
// COUNT adds the number of BVH nodes visited and triangles tested to the frame statistics.
constexpr uint32_t kRefillLanes = 16;
constexpr uint32_t kNoRay = 0xFFFFFFFFu;

// ALPHA: some mesh runs the alpha test (the opaque variant keeps six waves per SIMD);
// TEXA: some material has textures, so the test may read a texture's alpha.
template <bool COUNT, bool ALPHA, bool TEXA = false>
__global__ __launch_bounds__(kTraceBlock) __attribute__((amdgpu_waves_per_eu(TEXA ? 4 : 6))) void kExtend(SceneView sv, const float4* ro, const float4* rd, float4* hits,
                                                      uint32_t* counters, int countSlot, uint32_t* gstack,
                                                      uint32_t alphaSalt, uint32_t pathBase, float4 camO) {
    // camO.w != 0: the queue holds primary rays without stored origins; camO.xyz is theirs
    uint32_t nodeVisits = 0, triTests = 0;
    __shared__ uint32_t lds[kLdsStackWords];
    TraversalStack st;
    st.lds = lds + threadIdx.x;
    st.glob = gstack + (size_t)(blockIdx.x * kTraceBlock + threadIdx.x) * kGlobalStackWords;
    st.dropped = 0;
    st.sp = 0;
    const uint32_t n = counters[countSlot];
    if (n == 0u || sv.nodeCount == 0u) {
        // empty queue, or no geometry: every ray misses
        for (uint32_t i = blockIdx.x * kTraceBlock + threadIdx.x; i < n; i += gridDim.x * kTraceBlock)
            hits[i] = make_float4(rd[i].w, __uint_as_float(kNoHit), 0.0f, 0.0f);
        return;
    }
    // Wavefronts beyond one per 64 queued rays leave at once: every ray still
    // has a wavefront to claim it (each wavefront drains all partitions), and a
    // short queue does not pay kFetchParts contended claims in every resident
    // wavefront (about 100 us per launch with the full grid).
    if ((blockIdx.x * (kTraceBlock / 64u) + (threadIdx.x >> 6)) * 64ull >= n) return;
    const uint32_t lane = threadIdx.x & 63u;
    const uint64_t lanesBelow = (1ull << lane) - 1ull;
    const uint32_t partSize = (n + kFetchParts - 1u) / kFetchParts;
    uint32_t part = blockIdx.x % kFetchParts;
    uint32_t partsTried = 0;  // partitions found exhausted by this wavefront
    uint32_t rayId = kNoRay, alphaKey = 0u;
    TraversalRay r;
    ClosestState cs;
    cs.ref = 0;
    cs.steps = 0;
    cs.h.t = 0.0f;
    cs.h.tri = kNoHit;
    cs.h.u = cs.h.v = 0.0f;
    while (true) {
        const uint64_t idleMask = __ballot(rayId == kNoRay);
        const uint32_t idle = (uint32_t)__popcll(idleMask);
        if (partsTried < kFetchParts && (idle >= kRefillLanes || idle == 64u)) {
            // claim `idle` rays of the current partition; lanes beyond its end stay idle
            uint32_t base = 0;
            if (lane == 0u) base = atomicAdd(&counters[kCntFetchExt + part * kFetchStride], idle);
            base = (uint32_t)__shfl((int)base, 0);
            const uint32_t begin = part * partSize, end = min(n, begin + partSize);
            if (rayId == kNoRay) {
                const uint32_t k = begin + base + (uint32_t)__popcll(idleMask & lanesBelow);
                if (base < partSize && k < end) {
                    float4 a = camO.w != 0.0f ? make_float4(camO.x, camO.y, camO.z, __uint_as_float(k)) : ro[k];
                    float4 b = rd[k];
                    r = makeTraversalRay(xyz(a), xyz(b));
                    if (reachesSceneBox(sv, r, kRayTMin, b.w)) {
                        closestBegin(cs, st, b.w);
                        if (ALPHA) alphaKey = alphaRayKey(alphaSalt, pathBase + __float_as_uint(a.w));
                        rayId = k;
                    } else {  // misses every triangle: the lane is free again at once
                        hits[k] = make_float4(b.w, __uint_as_float(kNoHit), 0.0f, 0.0f);
                    }
                }
            }
            if (base + idle >= partSize) {  // wave-uniform: this partition is used up
                part = (part + 1u) % kFetchParts;
                partsTried++;
            }
        }
        if (rayId != kNoRay) {
            if (!closestStep<ALPHA, TEXA>(sv, r, kRayTMin, cs, st, alphaKey, COUNT ? &nodeVisits : nullptr,
                                    COUNT ? &triTests : nullptr)) {
                hits[rayId] = make_float4(cs.h.t, __uint_as_float(cs.h.tri), cs.h.u, cs.h.v);
                rayId = kNoRay;
            }
        } else if (partsTried >= kFetchParts) {
            // nothing left to claim; this lane is done, but the wavefront keeps
            // iterating until every lane has finished its ray
        }
        if (partsTried >= kFetchParts && __ballot(rayId != kNoRay) == 0ull) break;
    }
    if (st.dropped) atomicAdd(&counters[kCntDropped], st.dropped);
    if (COUNT) {
        unsigned long long* stats = reinterpret_cast<unsigned long long*>(counters + kCntStats);
        atomicAdd(&stats[2], (unsigned long long)nodeVisits);
        atomicAdd(&stats[3], (unsigned long long)triTests);
    }
}

DEV uint32_t classifyOccluder(const SceneView& sv, uint32_t tri) {
    if (tri == kNoHit) return kVisVisible;
    const Material& m = sv.materials[sv.meshes[sv.triMesh[tri]].materialIndex];
    return m.transmission > 0.0f ? kVisUnsupported : kVisOccluded;
}

// Any-hit queries of one shadow queue. Each record carries the contribution if
// the segment is visible (a) and if it is not (b); at most one record per path
// is in a queue, so the radiance update needs no atomics. Light rays follow the
// reference's classification (a transmissive first occluder clears prevNee,
// src/shaders/entry/shadow/closest_hit.slang:4-8); environment rays are visible
// or not.
template <int MODE>
__global__ __launch_bounds__(kTraceBlock) void kShadow(SceneView sv, PathView pv, ShadowView sh, uint32_t* counters,
                                                      int countSlot, int fetchSlot, uint32_t envPass,
                                                      uint32_t* gstack, uint32_t alphaSalt, uint32_t pathBase) {
    __shared__ uint32_t lds[kLdsStackWords];
    TraversalStack st;
    st.lds = lds + threadIdx.x;
    st.glob = gstack + (size_t)(blockIdx.x * kTraceBlock + threadIdx.x) * kGlobalStackWords;
    st.dropped = 0;
    const uint32_t n = counters[countSlot];
    if (n == 0u) return;
    // A wavefront claims batches of its block's partition only, and partition p
    // holds batches p, p + kFetchParts, ...: keep the wavefronts whose rank within
    // their partition is below that partition's batch count (rank 0 always), so
    // short queues do not pay a contended claim in every resident wavefront.
    const uint32_t rank = (blockIdx.x / kFetchParts) * (kTraceBlock / 64u) + (threadIdx.x >> 6);
    if ((uint64_t)rank * kFetchBatch * kFetchParts >= n) return;
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t part = blockIdx.x % kFetchParts;
    while (true) {
        const uint32_t base = claimBatch(counters + fetchSlot, part, n);
        if (base >= n) break;
        uint32_t i = base + lane;
        if (i >= n) continue;
        float4 o = sh.o[i], d = sh.d[i];
        uint32_t q = __float_as_uint(o.w);
        TraversalRay r = makeTraversalRay(xyz(o), xyz(d));
        uint32_t tri = traceAny(sv, r, kRayTMin, d.w, st, alphaRayKey(alphaSalt, pathBase + q));
        uint32_t vis = envPass ? (tri == kNoHit ? kVisVisible : kVisOccluded) : classifyOccluder(sv, tri);
        if (vis != kVisVisible && envPass) continue;  // b is zero for environment rays
        F4 c = ld4(vis == kVisVisible ? sh.a[i] : sh.b[i]);
        uint32_t flags = pv.flags[q];
        if (vis == kVisUnsupported) pv.flags[q] = flags & ~kFPrevNee;
        if (MODE == kModeRgb) {
            F4 rad = ld4(pv.rad[q]);
            pv.rad[q] = st4(F3{rad.x, rad.y, rad.z} + F3{c.x, c.y, c.z}, 0.0f);
        } else if (MODE == kModeSingle) {
            F4 rad = ld4(pv.rad[q]);
            rad.x += c.x;
            pv.rad[q] = st4(rad);
        } else if (flags & (kFHeroActive | kFPendingCollapse)) {
            pv.rad[q] = st4(ld4(pv.rad[q]) + c);
        } else {
            F4 rx = ld4(pv.radXYZ[q]);
            rx.w += c.x;
            pv.radXYZ[q] = st4(rx);
        }
    }
    if (st.dropped) atomicAdd(&counters[kCntDropped], st.dropped);
}

// The 11 inputs of the neural cache at a surface vertex, in [0, 1]: position in
// the scene box, outgoing direction and shading normal as (theta/pi, phi/2pi),
// roughness and albedo; f[2].w carries the path index.
DEV void nrcInputs(const FrameParams& fp, const Surface& sf, const BMat& mat, F3 d, uint32_t q, float4* f) {
    const float invPi = 0.31830988618379067f, inv2Pi = 0.15915494309189535f;
    F3 pn = F3{sat((sf.p.x - fp.nrcBoxMin[0]) * fp.nrcBoxInv[0]), sat((sf.p.y - fp.nrcBoxMin[1]) * fp.nrcBoxInv[1]),
               sat((sf.p.z - fp.nrcBoxMin[2]) * fp.nrcBoxInv[2])};
    F3 wo = -d, n = sf.basis.n;
    float dt = acosf(clampf(wo.z, -1.0f, 1.0f)) * invPi, dp = atan2f(wo.y, wo.x) * inv2Pi;
    float nt = acosf(clampf(n.z, -1.0f, 1.0f)) * invPi, np = atan2f(n.y, n.x) * inv2Pi;
    if (dp < 0.0f) dp += 1.0f;
    if (np < 0.0f) np += 1.0f;
    f[0] = make_float4(pn.x, pn.y, pn.z, dt);
    f[1] = make_float4(dp, nt, np, mat.roughness);
    f[2] = make_float4(mat.baseColor.x, mat.baseColor.y, mat.baseColor.z, __uint_as_float(q));
}

// Training targets of the cache: the radiance the rest of the path carried back
// to each recorded vertex, divided by the throughput that arrived there, less
// the vertex's own emission. In RGB mode that is the target. In hero spectral
// mode the same difference is taken per wavelength lane and the cache learns
// its CIE XYZ (the hero estimator's own conversion of the four lanes), so every
// target is a non-negative, unbiased estimate; records of paths whose hero
// wavelengths collapsed later no longer hold per-lane radiance and are dropped.
// Records whose throughput is too small to divide by are marked with w = -1.
__global__ __launch_bounds__(kBlock) void kNrcTargets(const float4* record, const uint32_t* count, uint32_t capacity,
                                                     const float4* rad, const uint32_t* flags, uint32_t hero,
                                                     float4* target) {
    const uint32_t r = blockIdx.x * kBlock + threadIdx.x;
    const uint32_t n = min(*count, capacity);
    if (r >= n) return;
    const float4* rec = record + (size_t)r * kNrcRecordWords;
    const uint32_t path = __float_as_uint(rec[2].w);
    const float4 L = rad[path], b = rec[3], R = rec[4], E = rec[5];
    if (hero) {
        const float lc[4] = {L.x, L.y, L.z, L.w}, bc[4] = {b.x, b.y, b.z, b.w}, rc[4] = {R.x, R.y, R.z, R.w},
                    ec[4] = {E.x, E.y, E.z, E.w};
        float sl[4];
        bool ok = (flags[path] & kFHeroActive) != 0u;
        for (int c = 0; c < 4; c++) {
            ok = ok && bc[c] > 1e-4f;
            sl[c] = ok ? fmaxf((lc[c] - rc[c]) / bc[c] - ec[c], 0.0f) : 0.0f;
        }
        const F3 xyz = ok ? toXYZ4(F4{sl[0], sl[1], sl[2], sl[3]}, ld4(rec[6]), f4(470.0f)) : f3(0.0f);
        target[r] = make_float4(fminf(xyz.x, 1000.0f), fminf(xyz.y, 1000.0f), fminf(xyz.z, 1000.0f), ok ? 0.0f : -1.0f);
        return;
    }
    const float lc[3] = {L.x, L.y, L.z}, bc[3] = {b.x, b.y, b.z}, rc[3] = {R.x, R.y, R.z}, ec[3] = {E.x, E.y, E.z};
    float t[3];
    bool ok = true;
    for (int c = 0; c < 3; c++) {
        ok = ok && bc[c] > 1e-4f;
        t[c] = ok ? fminf(fmaxf((lc[c] - rc[c]) / bc[c] - ec[c], 0.0f), 1000.0f) : 0.0f;
    }
    target[r] = make_float4(t[0], t[1], t[2], ok ? 0.0f : -1.0f);
}

// One path vertex for every ray of the current queue, at three waves per SIMD
// (at most 168 registers). Every variant must fit that budget without spilling
// registers to scratch: a hero build that spilled about 60 registers gave wrong
// NEE results while the same source at two waves matched the oracle bit for
// bit. What keeps the hero variant inside the budget: the rgb2spec resolution is
// a compile-time constant, medium coefficients and the scalar-wavelength state
// are read only where used, and shadow rays are queued before BSDF sampling.
// Materials with sheen take SHEEN variants at two waves per SIMD: the sheen
// layer does not fit the three-wave budget of the hero kernel.
// One path vertex: the ray (ro, rdir) of path q = bits of ro.w, at `depth`,
// and its hit h. Leaves the continuation ray in nextO / nextD with cont set, or
// cont false when the path ends here. Shadow and cache records are appended to
// their queues by whichever lanes of the wave reach the append.
template <int MODE, bool NRC, bool TEX, bool SHEEN>
DEV void shadeVertex(const FrameParams& fp, uint32_t depth, const SceneView& sv, const PathView& pv, float4 ro,
                     float4 rdir, Hit h, const ShadowView& sh, const ShadowView& she, uint32_t* counters,
                     const NrcView& nv, bool& cont, F3& nextO, F3& nextD) {
    bool shadowEmit = false, envEmit = false;
    F3 shO = f3(0.0f), shD = f3(0.0f), seO = f3(0.0f), seD = f3(0.0f);
    float shT = 0.0f;
    F4 shA = f4(0.0f), shB = f4(0.0f), seA = f4(0.0f);
    const uint32_t q = __float_as_uint(ro.w);
    do {
        const SceneData& sd = fp.sd;
        F3 o = xyz(ro), d = xyz(rdir);
        // depth 0 of a lean frame: the path state is the one kGenerate would have written
        const bool first = fp.leanStart != 0u && depth == 0u;
        PathStart ps{};
        if (first) ps = pathStart<MODE>(fp, q);
        // first read of the accumulated radiance in this vertex (0 at depth 0 of a lean frame)
        auto radIn = [&]() { return first ? f4(0.0f) : ld4(pv.rad[q]); };
        auto radXYZIn = [&]() { return first ? f4(0.0f) : ld4(pv.radXYZ[q]); };
        uint32_t flags = first ? (MODE == kModeHero ? kFHeroActive : 0u) : pv.flags[q];
        F4 lam = (MODE != kModeRgb) ? (first ? ps.lam : ld4(pv.lam[q])) : f4(0.0f);
        if (first) {  // what the finalisation and the shadow passes read of every path
            pv.rad[q] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            if (MODE == kModeHero) pv.radXYZ[q] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            if (MODE != kModeRgb) pv.lam[q] = st4(lam);
            pv.flags[q] = flags;
        }
        if (MODE == kModeHero && (flags & kFPendingCollapse)) {
            // collapse to the hero wavelength (src/shaders/integrator/path/spectral_hero/transport.slang:77-87),
            // deferred from the previous vertex until its shadow rays had added their contributions
            F4 rx = ld4(pv.radXYZ[q]);
            F3 xyzv = F3{rx.x, rx.y, rx.z} + toXYZ4(ld4(pv.rad[q]), lam, f4(470.0f));
            pv.radXYZ[q] = make_float4(xyzv.x, xyzv.y, xyzv.z, rx.w);
            pv.rad[q] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            flags &= ~kFPendingCollapse;
        }
        const bool absorbing = (flags & kFAbsorption) != 0u;
        const bool refractive = (flags & kFRefractive) != 0u;
        const bool heroActive = MODE == kModeHero && (flags & kFHeroActive) != 0u;
        F4 thr = first ? f4(1.0f) : ld4(pv.thr[q]);
        // prevBsdfPdf, scalar throughput, scalar wavelength, scalar 1/pdf; unused while
        // all four hero wavelengths are alive
        F4 misc = heroActive ? f4(0.0f) : (first ? F4{0.0f, 1.0f, 0.0f, 0.0f} : ld4(pv.misc[q]));

        if (h.tri == kNoHit) {
            if (!fp.neeOnly && !(refractive || absorbing)) {
                F3 env = environmentRadiance(sv, sd, d);
                // MIS against environment NEE at the previous vertex (power heuristic; hero balance)
                const bool prevEnv = (flags & kFPrevEnvNee) != 0u;
                float envPdf = prevEnv ? environmentPdf(sv, sd, d) : 0.0f;
                float w = 1.0f;
                if (prevEnv && !heroActive && envPdf > 0.0f && misc.x > 0.0f) w = powerHeuristic(misc.x, envPdf);
                if (MODE == kModeRgb) {
                    F4 rad = radIn();
                    pv.rad[q] = st4(F3{rad.x, rad.y, rad.z} + F3{thr.x, thr.y, thr.z} * (env * w), 0.0f);
                } else if (MODE == kModeSingle) {
                    F4 rad = radIn();
                    rad.x += thr.x * (spec1(fp.spec, env, lam.x) * w);
                    pv.rad[q] = st4(rad);
                } else if (heroActive) {
                    F4 tpp = first ? f4(1.0f) : ld4(pv.tpp[q]);
                    float hw = heroBalance(tpp);
                    if (prevEnv && envPdf > 0.0f) {
                        F4 ptpp = ld4(pv.ptpp[q]), pbtp = ld4(pv.pbtp[q]);
                        hw = spectralMis(ptpp * pbtp, ptpp * envPdf);
                    }
                    pv.rad[q] = st4(radIn() + thr * hw * spec4(fp.spec, env, lam));
                } else {
                    F4 rx = radXYZIn();
                    rx.w += misc.y * (spec1(fp.spec, env, misc.z) * w);
                    pv.radXYZ[q] = st4(rx);
                }
            }
            break;
        }

        // the medium coefficients exist only inside an absorbing medium; they are
        // re-read where needed rather than held in registers through the vertex
        auto mediumSigma = [&]() { return absorbing ? ld4(pv.sigma[q]) : f4(0.0f); };
        if (absorbing) {
            const F4 sigma = mediumSigma();
            if (MODE == kModeRgb) {
                F3 t = transmittance3(sigma, true, h.t);
                thr = F4{thr.x * t.x, thr.y * t.y, thr.z * t.z, thr.w};
            } else if (MODE == kModeSingle) {
                thr.x *= transmittance3(sigma, true, h.t).x;
            } else if (heroActive) {
                thr = thr * transmittance4(sigma, true, h.t);
            } else {
                misc.y *= transmittance3(sigma, true, h.t).x;
            }
        }
        if (MODE == kModeRgb ? !anyPos(F3{thr.x, thr.y, thr.z})
                             : (MODE == kModeSingle ? !(thr.x > 0.0f) : (heroActive ? !anyPos(thr) : !(misc.y > 0.0f))))
            break;

        Surface sf = reconstruct<TEX>(sv, o, d, h);
        BMat mat = TEX ? loadSurfaceMaterial(sv, sf) : loadMaterial(sv.materials, sf.materialIndex, sf.vertexColor);
        F3 wo = toLocal(-d, sf.basis);
        const float stateLambda = MODE == kModeRgb ? 0.0f : (MODE == kModeSingle ? lam.x : (heroActive ? lam.x : misc.z));
        BState st = makeState<SHEEN>(mat, wo, sf.frontFace, stateLambda, MODE == kModeRgb ? 0u : 1u);
        const bool curNee = !refractive;
        // neural cache: training paths record every vertex; image paths stop at the query vertex
        // (hero spectral: only while the four wavelengths are alive; a path whose
        // wavelengths collapsed is traced on without the cache)
        const bool nrcLanes = MODE != kModeHero || heroActive;
        const bool nrcTrain = NRC && q >= fp.nrcRenderPaths;
        const bool nrcQuery = NRC && !nrcTrain && depth == fp.nrcPrefix && nrcLanes;
        Rng rng = first ? ps.rng : loadRng(pv.rng[q], fp.rngPolicy);
        F4 tpp = heroActive ? (first ? f4(1.0f) : ld4(pv.tpp[q])) : f4(0.0f);

        // next-event estimation (src/shaders/light/direct/*_lighting.slang): the contribution is
        // evaluated here assuming visibility; the shadow kernel picks it or drops it
        F4 nee = f4(0.0f);
        bool neeTried = false;
        if (!fp.bsdfOnly && fp.neeEnabled && curNee && wo.z > 0.0f && !nrcQuery) {
            LightSurfaceSample ls = sampleLightSurface(sv, sd, sf.p, rng);
            if (ls.valid) {
                neeTried = true;
                F3 off = dot(ls.wi, sf.ng) >= 0.0f ? sf.ng : -sf.ng;
                shO = sf.p + off * kShadowOriginOffset;
                shD = ls.wi;
                shT = ls.shadowDistance;
                F3 wl = toLocal(ls.wi, sf.basis);
                if (!(isRefractive(mat) && wl.z <= 0.0f)) {
                    if (MODE == kModeRgb) {
                        Eval e = evalScalar<SHEEN>(fp.spec, st, wl);
                        if (e.pdf > 0.0f) {
                            float mis = powerHeuristic(ls.pdfSolidAngle, e.pdf);
                            F3 tr = transmittance3(mediumSigma(), absorbing, ls.shadowDistance);
                            F3 fc = e.value * fabsf(wl.z);
                            F3 r = tr * mis * fc * ls.emission / ls.pdfSolidAngle;
                            nee = F4{r.x, r.y, r.z, 0.0f};
                        }
                    } else if (heroActive) {
                        F4 btp;
                        F4 fc = evalHero<SHEEN>(fp.spec, st, wl, lam, btp) * fabsf(wl.z);
                        float mis = spectralMis(tpp * ls.pdfSolidAngle, tpp * btp);
                        if (mis > 0.0f) {
                            F4 tr = transmittance4(mediumSigma(), absorbing, ls.shadowDistance);
                            nee = tr * mis * fc * spec4(fp.spec, ls.emission, lam) / ls.pdfSolidAngle;
                        }
                    } else {
                        Eval e = evalScalar<SHEEN>(fp.spec, st, wl);
                        if (e.pdf > 0.0f) {
                            float sv1 = e.value.x * fabsf(wl.z) * spec1(fp.spec, ls.emission, stateLambda);
                            float mis = powerHeuristic(ls.pdfSolidAngle, e.pdf);
                            nee.x = mis * transmittance3(mediumSigma(), absorbing, ls.shadowDistance).x * sv1 / ls.pdfSolidAngle;
                        }
                    }
                }
            }
        }

        // environment NEE (an addition): drawn after the light NEE, traced by its own shadow pass
        bool envTried = false;
        if (!fp.bsdfOnly && fp.envNee && curNee && !absorbing && wo.z > 0.0f && !nrcQuery) {
            envTried = true;
            EnvSample es = sampleEnvironmentDirection(sv, sd, rng);
            if (es.valid) {
                F3 wl = toLocal(es.wi, sf.basis);
                if (!(isRefractive(mat) && wl.z <= 0.0f)) {
                    F3 le = environmentRadiance(sv, sd, es.wi);
                    F4 ne = f4(0.0f);
                    if (MODE == kModeRgb) {
                        Eval e = evalScalar<SHEEN>(fp.spec, st, wl);
                        if (e.pdf > 0.0f) {
                            float mis = powerHeuristic(es.pdf, e.pdf);
                            F3 r = (e.value * fabsf(wl.z)) * mis * le / es.pdf;
                            ne = F4{r.x, r.y, r.z, 0.0f};
                        }
                    } else if (heroActive) {
                        F4 btp;
                        F4 fc = evalHero<SHEEN>(fp.spec, st, wl, lam, btp) * fabsf(wl.z);
                        float mis = spectralMis(tpp * es.pdf, tpp * btp);
                        if (mis > 0.0f) ne = fc * mis * spec4(fp.spec, le, lam) / es.pdf;
                    } else {
                        Eval e = evalScalar<SHEEN>(fp.spec, st, wl);
                        if (e.pdf > 0.0f) {
                            float sv1 = e.value.x * fabsf(wl.z) * spec1(fp.spec, le, stateLambda);
                            ne.x = powerHeuristic(es.pdf, e.pdf) * sv1 / es.pdf;
                        }
                    }
                    if (anyPos(ne)) {
                        envEmit = true;
                        // start on the side the BSDF continuation ray for this direction would
                        // (+Ng for the shading reflection hemisphere), so both strategies see
                        // the same visibility and their MIS weights cover the same integrand
                        F3 off = wl.z > 0.0f ? sf.ng : -sf.ng;
                        seO = sf.p + off * kShadowOriginOffset;
                        seD = es.wi;
                        F4 thrE = (MODE == kModeHero && !heroActive) ? F4{misc.y, 0.0f, 0.0f, 0.0f} : thr;
                        seA = thrE * ne;
                    }
                }
            }
        }

        // emission with MIS (src/shaders/integrator/path/loop.slang:62-81, mis_weights.slang)
        F4 em = f4(0.0f);
        bool hasEm = false;
        if (!fp.neeOnly && anyPos(mat.emission)) {
            hasEm = true;
            const bool prevNee = (flags & kFPrevNee) != 0u;
            if (MODE == kModeHero && heroActive) {
                float w = heroBalance(tpp);
                if (fp.neeEnabled && prevNee && depth > 0u) {
                    float lp = areaToSolidAngle(sv.meshes[sf.mesh].lightPdfArea, sf.ng, d, h.t);
                    F4 ptpp = ld4(pv.ptpp[q]), pbtp = ld4(pv.pbtp[q]);
                    w = spectralMis(ptpp * pbtp, ptpp * lp);
                }
                em = spec4(fp.spec, mat.emission, lam) * w;
            } else {
                float w = 1.0f;
                if (fp.neeEnabled && prevNee && depth != 0u && misc.x > 0.0f) {
                    float lp = areaToSolidAngle(sv.meshes[sf.mesh].lightPdfArea, sf.ng, d, h.t);
                    w = lp <= 0.0f ? 1.0f : powerHeuristic(misc.x, lp);
                }
                if (MODE == kModeRgb)
                    em = F4{mat.emission.x * w, mat.emission.y * w, mat.emission.z * w, 0.0f};
                else
                    em.x = spec1(fp.spec, mat.emission, stateLambda) * w;
            }
        }

        if (NRC && nrcTrain && nrcLanes && wo.z > 0.0f) {
            const uint32_t slot = waveAppend(true, &counters[kCntNrcRecord]);
            if (slot < nv.recordCapacity) {
                float4* r = nv.record + (size_t)slot * kNrcRecordWords;
                nrcInputs(fp, sf, mat, d, q, r);
                if (MODE == kModeRgb) {
                    r[3] = make_float4(thr.x, thr.y, thr.z, 0.0f);
                    r[4] = st4(radIn());
                    r[5] = hasEm ? make_float4(em.x, em.y, em.z, 0.0f) : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
                    r[6] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
                } else {  // hero: four wavelength lanes and the wavelengths
                    r[3] = st4(thr);
                    r[4] = st4(radIn());
                    r[5] = hasEm ? st4(em) : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
                    r[6] = st4(lam);
                }
            }
        }

        // radiance += throughput * contribution, now or once visibility is known
        F4 thrUse = (MODE == kModeHero && !heroActive) ? F4{misc.y, 0.0f, 0.0f, 0.0f} : thr;
        if (neeTried && (fp.traceShadowAlways || anyPos(nee))) {
            shadowEmit = true;
            F4 cv = f4(0.0f) + nee;
            if (hasEm) cv = cv + em;
            F4 ch = f4(0.0f) + f4(0.0f);
            if (hasEm) ch = ch + em;
            shA = thrUse * cv;
            shB = thrUse * ch;
        } else if (hasEm || neeTried) {
            F4 cv = f4(0.0f);
            if (neeTried) cv = cv + f4(0.0f);
            if (hasEm) cv = cv + em;
            F4 add = thrUse * cv;
            if (MODE == kModeRgb) {
                F4 rad = radIn();
                pv.rad[q] = st4(F3{rad.x, rad.y, rad.z} + F3{add.x, add.y, add.z}, 0.0f);
            } else if (MODE == kModeSingle) {
                F4 rad = radIn();
                rad.x += add.x;
                pv.rad[q] = st4(rad);
            } else if (heroActive) {
                pv.rad[q] = st4(radIn() + add);
            } else {
                F4 rx = radXYZIn();
                rx.w += add.x;
                pv.radXYZ[q] = st4(rx);
            }
        }

        // The shadow payloads are final here: queue them now rather than at the end
        // so they are not held in registers through BSDF sampling (waveAppend works
        // on whichever lanes are still active).
        {
            const uint32_t sslot = waveAppend(shadowEmit, &counters[kCntShadow]);
            if (shadowEmit) {
                sh.o[sslot] = make_float4(shO.x, shO.y, shO.z, __uint_as_float(q));
                sh.d[sslot] = make_float4(shD.x, shD.y, shD.z, shT);
                sh.a[sslot] = st4(shA);
                sh.b[sslot] = st4(shB);
            }
            const uint32_t eslot = waveAppend(envEmit, &counters[kCntShadowEnv]);
            if (envEmit) {
                she.o[eslot] = make_float4(seO.x, seO.y, seO.z, __uint_as_float(q));
                she.d[eslot] = make_float4(seD.x, seD.y, seD.z, kRayTMax);
                she.a[eslot] = st4(seA);
            }
        }

        // the cache stands in for the rest of an image path from its query vertex on
        if (NRC && nrcQuery) {
            if (wo.z > 0.0f) {
                const uint32_t slot = waveAppend(true, &counters[kCntNrcQuery]);
                if (slot < nv.queryCapacity) {
                    float4* r = nv.query + (size_t)slot * kNrcQueryWords;
                    nrcInputs(fp, sf, mat, d, q, r);
                    if (MODE == kModeRgb) {
                        r[3] = make_float4(thr.x, thr.y, thr.z, 0.0f);
                        r[4] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
                    } else {
                        // hero: the cached XYZ radiance enters like the radiance of a
                        // source seen by all four lanes, each lane one quarter of the
                        // estimator's four-wavelength average
                        r[3] = st4(thr * 0.25f);
                        r[4] = st4(lam);
                    }
                }
            }
            break;
        }

        // the last vertex of the loop never traces its continuation
        if (depth + 1u >= sd.rrMaxDepth) break;

        // BSDF sampling (src/shaders/bsdf/sample_rgb.slang, sample_spectral.slang)
        if (wo.z <= 0.0f) break;
        F3 wl;
        uint32_t isT;
        if (!sampleLobe<SHEEN>(st, rng, wl, isT)) break;
        F3 wi = toWorld(wl, sf.basis);
        uint32_t newFlags = flags & ~(kFPrevNee | kFPendingCollapse | kFPrevEnvNee);
        if (heroActive) {
            F4 btp;
            F4 val = evalHero<SHEEN>(fp.spec, st, wl, lam, btp);
            float sp = btp.x;
            if (!(sp > 0.0f)) break;
            F4 weight = val * fabsf(wl.z) / sp;
            if (!anyPos(weight)) break;
            pv.ptpp[q] = st4(tpp);
            pv.pbtp[q] = st4(btp);
            tpp = tpp * btp;
            thr = thr * weight;
            if (!anyPos(thr)) break;
            if (isT != 0u && isRefractive(mat) && mat.abbeNumber > 0.0f) {
                // collapse to the hero wavelength (spectral_hero/transport.slang:77-87)
                misc.y = thr.x;
                thr = f4(0.0f);
                misc.z = lam.x;
                misc.w = 470.0f;
                misc.x = btp.x;
                newFlags &= ~kFHeroActive;
                if (shadowEmit || envEmit) {
                    newFlags |= kFPendingCollapse;
                } else {
                    F4 rx = ld4(pv.radXYZ[q]);  // this vertex may have added to both already
                    F3 xyzv = F3{rx.x, rx.y, rx.z} + toXYZ4(ld4(pv.rad[q]), lam, f4(470.0f));
                    pv.radXYZ[q] = make_float4(xyzv.x, xyzv.y, xyzv.z, rx.w);
                    pv.rad[q] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
                }
            }
        } else {
            Eval e = evalScalar<SHEEN>(fp.spec, st, wl);
            if (e.pdf <= 0.0f) break;
            F3 weight = e.value * fabsf(wl.z) / e.pdf;
            if (!(e.pdf > 0.0f && anyPos(weight))) break;
            if (MODE == kModeRgb) {
                F3 t = F3{thr.x, thr.y, thr.z} * weight;
                thr = F4{t.x, t.y, t.z, 0.0f};
                if (!anyPos(t)) break;
            } else if (MODE == kModeSingle) {
                thr.x *= weight.x;
                if (thr.x <= 0.0f) break;
            } else {
                misc.y *= weight.x;
                if (misc.y <= 0.0f) break;
            }
            misc.x = e.pdf;
        }
        const bool nowHero = MODE == kModeHero && (newFlags & kFHeroActive) != 0u;
        F4 sigma = mediumSigma();
        if (curNee && isT == 0u) newFlags |= kFPrevNee;
        if (envTried && isT == 0u) newFlags |= kFPrevEnvNee;
        if (nowHero)
            updateMedium4(fp.spec, mat, sf.frontFace, isT, lam, newFlags, sigma);
        else
            updateMedium(fp.spec, mat, sf.frontFace, isT, MODE == kModeRgb ? 0.0f : (MODE == kModeSingle ? lam.x : misc.z),
                         MODE == kModeRgb ? 0u : 1u, newFlags, sigma);

        // Russian roulette (rgb/transport.slang:56-62, spectral_hero/transport.slang:138-152)
        if (depth + 1u >= sd.rrMinDepth) {
            float p;
            if (MODE == kModeRgb)
                p = clampf(maxc(F3{thr.x, thr.y, thr.z}), kRRMin, kRRMax);
            else if (MODE == kModeSingle)
                p = clampf(thr.x, kRRMin, kRRMax);
            else
                p = nowHero ? clampf(maxc(thr), kRRMin, kRRMax) : clampf(misc.y, kRRMin, kRRMax);
            if (rnd(rng) > p) break;
            if (MODE == kModeRgb) {
                F3 t = F3{thr.x, thr.y, thr.z} / p;
                thr = F4{t.x, t.y, t.z, 0.0f};
            } else if (MODE == kModeSingle) {
                thr.x /= p;
            } else if (nowHero) {
                tpp = tpp * F4{p, p, p, p};
                thr = thr / p;
            } else {
                misc.y /= p;
            }
        }
        // src/shaders/integrator/path/loop.slang:113-120
        F3 off = isT ? -sf.ng : sf.ng;
        nextO = sf.p + off * kShadowOriginOffset;
        nextD = wi;
        cont = true;
        pv.rng[q] = storeRng(rng);
        pv.thr[q] = st4(thr);
        if (!nowHero) pv.misc[q] = st4(misc);
        pv.flags[q] = newFlags;
        if (newFlags & kFAbsorption) pv.sigma[q] = st4(sigma);
        if (MODE == kModeHero) pv.tpp[q] = st4(tpp);
    } while (false);
}

template <int MODE, bool NRC = false, bool TEX = false, bool SHEEN = false>
__global__ __launch_bounds__(kBlock, SHEEN ? 2 : 3) void kShade(FrameParams fp, SceneView sv, PathView pv, QueueView in,
                                                 const float4* hits, QueueView out, ShadowView sh, ShadowView she,
                                                 uint32_t* counters, int inSlot, int outSlot, NrcView nv,
                                                 const uint32_t* __restrict__ order) {
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    const uint32_t n = counters[inSlot];
    bool cont = false;
    F3 nextO = f3(0.0f), nextD = f3(0.0f);
    uint32_t q = 0;
    if (i < n) {
        const uint32_t j = order ? order[i] : i;  // shading order (sorted by material class) or queue order
        // depth 0 of a compact frame: the ray's origin is the camera's, its path index its queue slot
        F3 co = f3(0.0f);
        if (fp.compactPrimary && fp.depth == 0u) co = cameraOrigin(fp.sd);
        const float4 ro = fp.compactPrimary && fp.depth == 0u ? make_float4(co.x, co.y, co.z, __uint_as_float(j)) : in.o[j];
        const float4 rdir = in.d[j], hh = hits[j];
        q = __float_as_uint(ro.w);
        shadeVertex<MODE, NRC, TEX, SHEEN>(fp, fp.depth, sv, pv, ro, rdir, Hit{hh.x, __float_as_uint(hh.y), hh.z, hh.w},
                                           sh, she, counters, nv, cont, nextO, nextD);
    }
    uint32_t slot = waveAppend(cont, &counters[outSlot]);
    if (cont) {
        out.o[slot] = make_float4(nextO.x, nextO.y, nextO.z, __uint_as_float(q));
        out.d[slot] = make_float4(nextD.x, nextD.y, nextD.z, kRayTMax);
    }
}

// The last depths of a chunk in one launch, once the queues have become short:
// every lane takes a path of the queue at depth startDepth and carries it to
// its end, tracing and shading vertex after vertex (the arithmetic of kExtend
// and kShade, so the same image), instead of one launch per depth that waits
// for the slowest ray of that depth. A lane that has its hit waits until
// kTailShadeLanes lanes of the wave have theirs, or no lane is still tracing,
// and they shade together; idle lanes claim new paths in batches as in
// kExtend. Used when no shadow rays, cache, features, material sort, alpha
// test, material textures or sheen are active (the shading variant without
// them), at two waves per SIMD: the shading body inside the lane loop needs
// more registers than the three-wave budget of kShade.
constexpr uint32_t kTailShadeLanes = 48;  // of 32, 48, 56 and 64 lanes, 48 and 56 measured fastest
template <int MODE>
__global__ __launch_bounds__(kTraceBlock, 2) void kTail(FrameParams fp, SceneView sv, PathView pv, QueueView in,
                                                        uint32_t* counters, int countSlot, uint32_t startDepth,
                                                        uint32_t* gstack) {
    __shared__ uint32_t lds[kLdsStackWords];
    TraversalStack st;
    st.lds = lds + threadIdx.x;
    st.glob = gstack + (size_t)(blockIdx.x * kTraceBlock + threadIdx.x) * kGlobalStackWords;
    st.dropped = 0;
    st.sp = 0;
    const uint32_t n = counters[countSlot];
    // wavefronts beyond one per 64 queued paths leave at once
    if ((blockIdx.x * (kTraceBlock / 64u) + (threadIdx.x >> 6)) * 64ull >= n) return;
    constexpr uint32_t kIdle = 0u, kTrace = 1u, kReady = 2u;
    const uint32_t lane = threadIdx.x & 63u;
    const uint64_t lanesBelow = (1ull << lane) - 1ull;
    uint32_t state = kIdle, depth = 0u, traced = 0u;
    bool exhausted = false;  // wave-uniform: the queue has been claimed to its end
    float4 ro = make_float4(0.0f, 0.0f, 0.0f, 0.0f), rd = ro;
    TraversalRay r;
    ClosestState cs;
    cs.ref = 0;
    cs.steps = 0;
    cs.h.t = 0.0f;
    cs.h.tri = kNoHit;
    cs.h.u = cs.h.v = 0.0f;
    // start the traversal of the lane's ray (ro, rd): a ray that misses the
    // scene's box (or a scene without geometry) has its miss at once
    auto beginRay = [&]() {
        r = makeTraversalRay(xyz(ro), xyz(rd));
        if (sv.nodeCount > 0u && reachesSceneBox(sv, r, kRayTMin, rd.w)) {
            closestBegin(cs, st, rd.w);
            state = kTrace;
        } else {
            cs.h.t = rd.w;
            cs.h.tri = kNoHit;
            cs.h.u = cs.h.v = 0.0f;
            state = kReady;
        }
    };
    while (true) {
        const uint64_t idleMask = __ballot(state == kIdle);
        const uint32_t idle = (uint32_t)__popcll(idleMask);
        if (!exhausted && (idle >= kRefillLanes || idle == 64u)) {
            uint32_t base = 0;
            if (lane == 0u) base = atomicAdd(&counters[kCntTailFetch], idle);
            base = (uint32_t)__shfl((int)base, 0);
            if (base + idle >= n) exhausted = true;
            if (state == kIdle) {
                const uint32_t k = base + (uint32_t)__popcll(idleMask & lanesBelow);
                if (k < n) {
                    ro = in.o[k];
                    rd = in.d[k];
                    depth = startDepth;
                    beginRay();
                }
            }
        }
        if (state == kTrace && !closestStep<false, false>(sv, r, kRayTMin, cs, st, 0u)) state = kReady;
        const uint32_t ready = (uint32_t)__popcll(__ballot(state == kReady));
        if (ready >= kTailShadeLanes || (ready > 0u && __ballot(state == kTrace) == 0ull)) {
            if (state == kReady) {
                bool cont = false;
                F3 nextO = f3(0.0f), nextD = f3(0.0f);
                shadeVertex<MODE, false, false, false>(fp, depth, sv, pv, ro, rd, cs.h, ShadowView{}, ShadowView{},
                                                       counters, NrcView{}, cont, nextO, nextD);
                if (cont) {
                    ro = make_float4(nextO.x, nextO.y, nextO.z, ro.w);
                    rd = make_float4(nextD.x, nextD.y, nextD.z, kRayTMax);
                    depth++;
                    traced++;
                    beginRay();
                } else {
                    state = kIdle;
                }
            }
        }
        if (exhausted && __ballot(state != kIdle) == 0ull) break;
    }
    if (st.dropped) atomicAdd(&counters[kCntDropped], st.dropped);
    // the continuation rays (the queue's own rays were counted when the depth began)
    for (int off = 32; off > 0; off >>= 1) traced += (uint32_t)__shfl_down((int)traced, off);
    if (lane == 0u && traced) {
        unsigned long long* stats = reinterpret_cast<unsigned long long*>(counters + kCntStats);
        atomicAdd(&stats[0], (unsigned long long)traced);
        atomicAdd(&stats[4], (unsigned long long)traced);
    }
}

// The radiance path i delivers to its pixel (RGB, or XYZ in the spectral modes).
template <int MODE>
DEV F3 pathRadiance(const PathView& pv, uint32_t i) {
    F4 rad = ld4(pv.rad[i]);
    F3 out;
    if (MODE == kModeRgb) {
        out = F3{rad.x, rad.y, rad.z};
    } else if (MODE == kModeSingle) {
        F4 l = ld4(pv.lam[i]);
        out = toXYZ1(rad.x, l.x, l.y);
    } else {
        F4 rx = ld4(pv.radXYZ[i]);
        out = F3{rx.x, rx.y, rx.z};
        if (pv.flags[i] & kFHeroActive) {
            out = out + toXYZ4(rad, ld4(pv.lam[i]), f4(470.0f));
        } else {
            F4 m = ld4(pv.misc[i]);
            out = out + toXYZ1(rx.w, m.z, m.w);
        }
    }
    return out;
}

template <int MODE>
__global__ __launch_bounds__(kBlock) void kFinalize(uint32_t count, PathView pv, float4* sampleRad) {
    uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= count) return;
    F3 out = pathRadiance<MODE>(pv, i);
    sampleRad[i] = make_float4(out.x, out.y, out.z, 0.0f);
}

// kFinalize and kReduce in one pass: frameSum[k] += the radiance of every sample
// of pixel k inside [pathBegin, pathBegin+count), in sample order. Neighbouring
// threads are neighbouring pixels, so every sample's loads are coalesced.
template <int MODE>
__global__ __launch_bounds__(kBlock) void kFinalizeReduce(uint32_t numPixels, uint64_t pathBegin, uint32_t count,
                                                          PathView pv, float4* frameSum) {
    uint32_t k = blockIdx.x * kBlock + threadIdx.x;
    if (k >= numPixels) return;
    const uint64_t end = pathBegin + count;
    if (end <= k) return;  // no sample of this pixel in the chunk
    // samples s with pathBegin <= s * numPixels + k < end, counted before the
    // loop so that its loads can be issued ahead of the additions
    const uint64_t sBegin = pathBegin > k ? (pathBegin - k + numPixels - 1u) / numPixels : 0u;
    const uint64_t sEnd = (end - 1u - k) / numPixels + 1u;
    F4 acc = ld4(frameSum[k]);
    uint32_t i = (uint32_t)(sBegin * numPixels + k - pathBegin);
#pragma unroll 4
    for (uint64_t s = sBegin; s < sEnd; s++, i += numPixels) {
        F3 v = pathRadiance<MODE>(pv, i);
        acc = F4{acc.x + v.x, acc.y + v.y, acc.z + v.z, 0.0f};
    }
    frameSum[k] = st4(acc);
}

// frameSum[k] += radiance of every sample of pixel k inside [pathBegin, pathBegin+count), in sample order
__global__ __launch_bounds__(kBlock) void kReduce(uint32_t numPixels, uint64_t pathBegin, uint32_t count,
                                                  const float4* sampleRad, float4* frameSum) {
    uint32_t k = blockIdx.x * kBlock + threadIdx.x;
    if (k >= numPixels) return;
    uint64_t end = pathBegin + count;
    uint64_t s0 = pathBegin / numPixels;
    F4 acc = ld4(frameSum[k]);
    for (uint64_t s = s0;; s++) {
        uint64_t p = s * numPixels + k;
        if (p >= end) break;
        if (p < pathBegin) continue;
        float4 v = sampleRad[p - pathBegin];
        acc = F4{acc.x + v.x, acc.y + v.y, acc.z + v.z, 0.0f};
    }
    frameSum[k] = st4(acc);
}

DEV float srgbEncode(float v) {
    if (v <= 0.0031308f) return 12.92f * v;
    return 1.055f * powf(v, 1.0f / 2.4f) - 0.055f;
}

// src/shaders/integrator/path/writeback.slang:9-123, film/tonemap.slang:9-31
__global__ __launch_bounds__(kBlock) void kAccumulate(SceneData sd, uint32_t width, uint32_t height, uint32_t vw,
                                                      uint32_t vh, const float4* frameSum, float4* accum,
                                                      ushort4* output, uint32_t spectralOutput) {
    uint32_t idx = blockIdx.x * kBlock + threadIdx.x;
    if (idx >= width * height) return;
    uint32_t x = idx % width, y = idx / width;
    // vw, vh: viewport extent clamped to the accumulation image, so k < vw * vh
    bool inside = vw > 0 && vh > 0 && x >= sd.viewportRect[0] && y >= sd.viewportRect[1] &&
                  x < sd.viewportRect[0] + vw && y < sd.viewportRect[1] + vh;
    if (!inside) {
        accum[idx] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        output[idx] = make_ushort4(0, 0, 0, 0);
        return;
    }
    uint32_t k = (y - sd.viewportRect[1]) * vw + (x - sd.viewportRect[0]);
    uint32_t spp = sd.samplesPerPixel > 1u ? sd.samplesPerPixel : 1u;
    F4 fs = ld4(frameSum[k]);
    F3 frame = F3{fs.x, fs.y, fs.z} / (float)spp;
    float4 prev = accum[idx];
    float pw = (float)(uint32_t)(prev.w + 0.5f), cw = (float)spp, tw = pw + cw;
    F3 acc = tw <= 0.0f ? f3(0.0f) : (F3{prev.x, prev.y, prev.z} * pw + frame * cw) / tw;
    accum[idx] = make_float4(acc.x, acc.y, acc.z, tw);
    F3 c = spectralOutput ? xyzToSrgb(acc) : acc;
    c = maxs(c * sd.exposure, 0.0f);
    if (VKRT_SCENE_TONE_MAPPING_MODE_FROM_RENDER_SETTINGS(sd.packedRenderSettings) == VKRT_TONE_MAPPING_MODE_ACES)
        c = (c * (c * 2.51f + f3(0.03f))) / (c * (c * 2.43f + f3(0.59f)) + f3(0.14f));
    c = sat(c);
    float e[3] = {srgbEncode(c.x), srgbEncode(c.y), srgbEncode(c.z)};
    unsigned short o[3];
    for (int j = 0; j < 3; j++) o[j] = (unsigned short)floorf(sat(e[j]) * 65535.0f + 0.5f);
    output[idx] = make_ushort4(o[0], o[1], o[2], 65535);
}

// ---------------------------------------------------- features, debug --
// Denoiser feature images (src/shaders/integrator/path/loop.slang:83-111,
// bsdf/principled/denoiser.slang) and the debug views (debug.slang,
// utility/debug.slang, writeback.slang:25-58). They run as passes of their
// own, only while feature images or a debug view are asked for, so the
// shading kernels keep their register budget.
struct FeatureView {
    uint32_t* state;  // per path: bit 0 resolved, bit 1 followed a specular hit at depth 0
    float4* albedo;   // per path: resolved albedo, w = resolve depth + 1
    float4* normal;   // per path: resolved shading normal
    float4* debug0;   // per pixel: primary-surface debug value of sample 0, w = 1 once set
    uint32_t* bounce0;  // per pixel: bounce count of sample 0
};

DEV bool isPrimarySurfaceDebug(uint32_t mode) {
    return mode == VKRT_DEBUG_MODE_NORMALS || mode == VKRT_DEBUG_MODE_DEPTH ||
           (mode >= VKRT_DEBUG_MODE_BASE_COLOR_MAP && mode <= VKRT_DEBUG_MODE_EMISSIVE_MAP);
}

// denoiser.slang:7-22
DEV bool denoiserFollowsSpecular(const BMat& m, uint32_t frontFace) {
    const float eps = 1e-3f;
    const float diffuseW = maxc(diffuseColor(m)) * (1.0f - m.transmission);
    const float subsurfaceW = diffuseW * m.subsurface;
    const float sheenW = maxc(m.sheenColor) * (1.0f - m.transmission);
    const float coatW = frontFace != 0u ? m.clearcoat : 0.0f;
    if (diffuseW > eps || subsurfaceW > eps || sheenW > eps || coatW > eps) return false;
    if (m.transmission > eps && dielectricIdentity(m)) return true;
    if (m.roughness > 0.02f) return false;
    return m.transmission > eps || m.metallic >= 1.0f - eps || maxc(specularF0(m)) > eps;
}

// denoiser.slang:24-46
DEV F3 denoiserAlbedo(const BMat& m) {
    const float nonMetal = 1.0f - m.metallic;
    const F3 dc = diffuseColor(m), f0 = specularF0(m);
    const F3 metalF0 = hasConductor(m) ? metalFresnel(m, 1.0f) : sat(m.baseColor);
    const float base = nonMetal * (1.0f - m.transmission);
    const float diffuseW = base * (1.0f - m.subsurface), subsurfaceW = base * m.subsurface;
    const float specularW = base * sat(lum(specularF0(m)));
    const float transmissionW = nonMetal * m.transmission, metallicW = m.metallic;
    const float sheenW = maxc(m.sheenColor) * base;
    const float total = diffuseW + subsurfaceW + specularW + transmissionW + metallicW + sheenW;
    if (total <= 0.0f) return sat(dc);
    const F3 a = dc * diffuseW + dc * subsurfaceW + (f0 * specularW + f3(transmissionW)) +
                 (metalF0 * metallicW + m.sheenColor * sheenW);
    return sat(a / total);
}

// debug.slang:30-66, utility/debug.slang:13-34
template <bool TEX>
DEV F3 primarySurfaceDebugValue(const SceneView& sv, uint32_t mode, const Surface& sf, float t) {
    if (mode == VKRT_DEBUG_MODE_NORMALS) return sf.basis.n * 0.5f + f3(0.5f);
    if (mode == VKRT_DEBUG_MODE_DEPTH) return f3(1.0f / (1.0f + t));
    const uint32_t slot = mode == VKRT_DEBUG_MODE_BASE_COLOR_MAP  ? VKRT_MATERIAL_TEXTURE_SLOT_BASE_COLOR
                          : mode == VKRT_DEBUG_MODE_NORMAL_MAP   ? VKRT_MATERIAL_TEXTURE_SLOT_NORMAL
                          : mode == VKRT_DEBUG_MODE_EMISSIVE_MAP ? VKRT_MATERIAL_TEXTURE_SLOT_EMISSIVE
                                                                 : VKRT_MATERIAL_TEXTURE_SLOT_METALLIC_ROUGHNESS;
    const float fb = slot == VKRT_MATERIAL_TEXTURE_SLOT_NORMAL ? 0.5f : 1.0f;
    F4 v = F4{fb, fb, 1.0f, 1.0f};
    if constexpr (TEX) v = sampleMaterialSlot(sv, sv.materials[sf.materialIndex], slot, sf.tc, v);
    if (mode == VKRT_DEBUG_MODE_METALLIC_MAP) return f3(v.z);
    if (mode == VKRT_DEBUG_MODE_ROUGHNESS_MAP) return f3(v.y);
    return F3{v.x, v.y, v.z};
}

__global__ void kFeatInit(uint32_t* state, uint32_t count) {
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i < count) state[i] = 0u;
}

// One pass per depth over the hits kShade is about to shade: the path's
// state up to the medium test repeats kShade's, then sample 0 records the
// primary-surface debug value, every path resolves its features at its first
// vertex that is not a specular hit to follow, and sample 0 counts bounces.
template <int MODE, bool TEX>
__global__ __launch_bounds__(kBlock) void kFeatures(FrameParams fp, SceneView sv, PathView pv, QueueView in,
                                                    const float4* hits, const uint32_t* counters, int inSlot,
                                                    FeatureView fv) {
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= counters[inSlot]) return;
    const float4 oq = in.o[i], dq = in.d[i], hh = hits[i];
    const uint32_t q = __float_as_uint(oq.w);
    if (fp.nrc && q >= fp.nrcRenderPaths) return;  // training paths of the neural cache
    const Hit h{hh.x, __float_as_uint(hh.y), hh.z, hh.w};
    if (h.tri == kNoHit) return;
    const bool first = fp.leanStart != 0u && fp.depth == 0u;
    const uint32_t flags = first ? (MODE == kModeHero ? kFHeroActive : 0u) : pv.flags[q];
    const bool heroActive = MODE == kModeHero && (flags & kFHeroActive) != 0u;
    F4 thr = first ? f4(1.0f) : ld4(pv.thr[q]);
    F4 misc = heroActive ? f4(0.0f) : (first ? F4{0.0f, 1.0f, 0.0f, 0.0f} : ld4(pv.misc[q]));
    if (flags & kFAbsorption) {
        const F4 sigma = ld4(pv.sigma[q]);
        if (MODE == kModeRgb) {
            F3 t = transmittance3(sigma, true, h.t);
            thr = F4{thr.x * t.x, thr.y * t.y, thr.z * t.z, thr.w};
        } else if (MODE == kModeSingle) {
            thr.x *= transmittance3(sigma, true, h.t).x;
        } else if (heroActive) {
            thr = thr * transmittance4(sigma, true, h.t);
        } else {
            misc.y *= transmittance3(sigma, true, h.t).x;
        }
    }
    if (MODE == kModeRgb ? !anyPos(F3{thr.x, thr.y, thr.z})
                         : (MODE == kModeSingle ? !(thr.x > 0.0f) : (heroActive ? !anyPos(thr) : !(misc.y > 0.0f))))
        return;
    const F3 o = F3{oq.x, oq.y, oq.z}, d = F3{dq.x, dq.y, dq.z};
    const Surface sf = reconstruct<TEX>(sv, o, d, h);
    const uint64_t p = fp.pathBegin + q;
    const bool sample0 = p < fp.numPixels;
    const uint32_t dm = fp.sd.debugMode;
    if (sample0 && fp.depth == 0u && isPrimarySurfaceDebug(dm)) {
        const F3 v = primarySurfaceDebugValue<TEX>(sv, dm, sf, h.t);
        fv.debug0[p] = make_float4(v.x, v.y, v.z, 1.0f);
        return;
    }
    const BMat mat = TEX ? loadSurfaceMaterial(sv, sf) : loadMaterial(sv.materials, sf.materialIndex, sf.vertexColor);
    const uint32_t st = fv.state[q];
    uint32_t next = st;
    const bool follow = denoiserFollowsSpecular(mat, sf.frontFace);
    if (fp.depth == 0u && follow) next |= 2u;
    if (!(st & 1u) && !follow) {
        const F3 a = denoiserAlbedo(mat);
        fv.albedo[q] = make_float4(a.x, a.y, a.z, (float)(fp.depth + 1u));
        fv.normal[q] = make_float4(sf.basis.n.x, sf.basis.n.y, sf.basis.n.z, 0.0f);
        next |= 1u;
    }
    if (next != st) fv.state[q] = next;
    if (sample0) fv.bounce0[p] = fp.depth + 1u;
}

// frameFeat[3k..3k+2] += the features of every sample of pixel k in the chunk
__global__ __launch_bounds__(kBlock) void kFeatReduce(uint32_t numPixels, uint64_t pathBegin, uint32_t count,
                                                      FeatureView fv, float4* frameFeat) {
    const uint32_t k = blockIdx.x * kBlock + threadIdx.x;
    if (k >= numPixels) return;
    const uint64_t end = pathBegin + count;
    F4 a = ld4(frameFeat[3u * k]), n = ld4(frameFeat[3u * k + 1u]), f = ld4(frameFeat[3u * k + 2u]);
    for (uint64_t s = pathBegin / numPixels;; s++) {
        const uint64_t p = s * numPixels + k;
        if (p >= end) break;
        if (p < pathBegin) continue;
        const uint32_t q = (uint32_t)(p - pathBegin), st = fv.state[q];
        if (st & 1u) {
            const F4 pa = ld4(fv.albedo[q]), pn = ld4(fv.normal[q]);
            a = F4{a.x + pa.x, a.y + pa.y, a.z + pa.z, a.w + 1.0f};
            n = F4{n.x + pn.x, n.y + pn.y, n.z + pn.z, n.w + pa.w};
        }
        if (st & 2u) f.x += 1.0f;
    }
    frameFeat[3u * k] = st4(a);
    frameFeat[3u * k + 1u] = st4(n);
    frameFeat[3u * k + 2u] = st4(f);
}

// writeback.slang:68-72
DEV F3 blendAccumulated(F3 prev, float prevW, F3 cur, float curW) {
    const float tw = prevW + curW;
    if (tw <= 0.0f) return f3(0.0f);
    return (prev * prevW + cur * curW) / tw;
}

DEV ushort4 encodeDisplay16(F3 c) {
    c = sat(c);
    const float e[3] = {srgbEncode(c.x), srgbEncode(c.y), srgbEncode(c.z)};
    unsigned short o[3];
    for (int j = 0; j < 3; j++) o[j] = (unsigned short)floorf(sat(e[j]) * 65535.0f + 0.5f);
    return make_ushort4(o[0], o[1], o[2], 65535);
}

// kAccumulate with the feature images and the debug views
// (src/shaders/entry/path/raygen_*.slang:17-34, writeback.slang:9-123)
__global__ __launch_bounds__(kBlock) void kAccumulateFeatures(SceneData sd, uint32_t width, uint32_t height,
                                                              uint32_t vw, uint32_t vh, const float4* frameSum,
                                                              const float4* frameFeat, FeatureView fv, float4* accum,
                                                              float4* albedoImg, float4* normalImg, ushort4* output,
                                                              uint32_t spectralOutput, uint32_t haveFeatures) {
    const uint32_t idx = blockIdx.x * kBlock + threadIdx.x;
    if (idx >= width * height) return;
    const uint32_t x = idx % width, y = idx / width;
    const bool inside = vw > 0 && vh > 0 && x >= sd.viewportRect[0] && y >= sd.viewportRect[1] &&
                        x < sd.viewportRect[0] + vw && y < sd.viewportRect[1] + vh;
    if (!inside) {
        accum[idx] = albedoImg[idx] = normalImg[idx] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        output[idx] = make_ushort4(0, 0, 0, 0);
        return;
    }
    const uint32_t k = (y - sd.viewportRect[1]) * vw + (x - sd.viewportRect[0]);
    const uint32_t spp = sd.samplesPerPixel > 1u ? sd.samplesPerPixel : 1u;
    const uint32_t dm = sd.debugMode;
    bool early = false;
    F3 rad = f3(0.0f), albedo = f3(0.0f), normal = f3(0.0f);
    float weight = 0.0f;
    if (dm == VKRT_DEBUG_MODE_SELECTION_MASK) {
        rad = f3(0.05f);  // debug.slang:10-19 with no mesh selected
        early = true;
    } else if (isPrimarySurfaceDebug(dm) && fv.debug0[k].w > 0.0f) {
        const float4 v = fv.debug0[k];
        rad = F3{v.x, v.y, v.z};
        early = true;
    } else if (dm == VKRT_DEBUG_MODE_BOUNCE_COUNT) {
        const float maxD = sd.rrMaxDepth > 1u ? (float)sd.rrMaxDepth : 1.0f;
        const float t = clampf((float)fv.bounce0[k] / maxD, 0.0f, 1.0f);
        rad = F3{t, 1.0f - t, 0.0f};
        early = true;
    }
    if (!early) {
        const F4 fs = ld4(frameSum[k]);
        rad = F3{fs.x, fs.y, fs.z} / (float)spp;
        float depthSum = 0.0f, follow = 0.0f;
        if (haveFeatures) {
            const F4 a = ld4(frameFeat[3u * k]), n = ld4(frameFeat[3u * k + 1u]);
            follow = frameFeat[3u * k + 2u].x;
            weight = a.w;
            depthSum = n.w;
            if (weight > 0.0f) {
                albedo = F3{a.x, a.y, a.z} / weight;
                normal = F3{n.x, n.y, n.z} / weight;
            }
        }
        switch (dm) {
            case VKRT_DEBUG_MODE_DENOISER_ALBEDO: rad = albedo; early = true; break;
            case VKRT_DEBUG_MODE_DENOISER_NORMAL:
                rad = weight > 0.0f ? normal * 0.5f + f3(0.5f) : f3(0.0f);
                early = true;
                break;
            case VKRT_DEBUG_MODE_DENOISER_FEATURE_VALIDITY: rad = f3(sat(weight / (float)spp)); early = true; break;
            case VKRT_DEBUG_MODE_DENOISER_FEATURE_DEPTH: {
                float dd = 0.0f;
                if (weight > 0.0f) {
                    const float maxD = sd.rrMaxDepth > 1u ? (float)(sd.rrMaxDepth - 1u) : 1.0f;
                    dd = sat((depthSum / weight - 1.0f) / maxD);
                }
                rad = f3(dd);
                early = true;
                break;
            }
            case VKRT_DEBUG_MODE_DENOISER_FOLLOW_SPECULAR: {
                const float t = sat(follow / (float)spp);
                rad = F3{lerpf(0.05f, 1.0f, t), lerpf(0.05f, 0.6f, t), lerpf(0.05f, 0.0f, t)};
                early = true;
                break;
            }
            default: break;
        }
    }
    if (early) {  // writeback.slang:59-66: the debug value, no sample count, no features
        accum[idx] = make_float4(rad.x, rad.y, rad.z, 0.0f);
        albedoImg[idx] = normalImg[idx] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        output[idx] = encodeDisplay16(rad);
        return;
    }
    const float4 prev = accum[idx];
    const float pw = (float)(uint32_t)(prev.w + 0.5f), cw = (float)spp, tw = pw + cw;
    const F3 acc = blendAccumulated(F3{prev.x, prev.y, prev.z}, pw, rad, cw);
    accum[idx] = make_float4(acc.x, acc.y, acc.z, tw);
    const float4 pa = albedoImg[idx], pn = normalImg[idx];
    const F3 a = blendAccumulated(F3{pa.x, pa.y, pa.z}, pa.w, albedo, weight);
    const F3 n = blendAccumulated(F3{pn.x, pn.y, pn.z}, pn.w, normal, weight);
    albedoImg[idx] = make_float4(a.x, a.y, a.z, pa.w + weight);
    normalImg[idx] = make_float4(n.x, n.y, n.z, pn.w + weight);
    F3 c = spectralOutput ? xyzToSrgb(acc) : acc;
    c = maxs(c * sd.exposure, 0.0f);
    if (VKRT_SCENE_TONE_MAPPING_MODE_FROM_RENDER_SETTINGS(sd.packedRenderSettings) == VKRT_TONE_MAPPING_MODE_ACES)
        c = (c * (c * 2.51f + f3(0.03f))) / (c * (c * 2.43f + f3(0.59f)) + f3(0.14f));
    output[idx] = encodeDisplay16(c);
}

__global__ void kFill4(float4* p, size_t n, float4 v) {
    size_t i = (size_t)blockIdx.x * kBlock + threadIdx.x;
    if (i < n) p[i] = v;
}

__global__ __launch_bounds__(kTraceBlock) void kTraceQuery(SceneView sv, const float4* rays, uint32_t n,
                                                          float4* out, uint32_t* gstack, uint32_t anyHit) {
    __shared__ uint32_t lds[kLdsStackWords];
    TraversalStack st;
    st.lds = lds + threadIdx.x;
    st.glob = gstack + (size_t)(blockIdx.x * kTraceBlock + threadIdx.x) * kGlobalStackWords;
    st.dropped = 0;
    for (uint32_t i = blockIdx.x * kTraceBlock + threadIdx.x; i < n; i += gridDim.x * kTraceBlock) {
        float4 a = rays[2 * i], b = rays[2 * i + 1];
        TraversalRay r = makeTraversalRay(xyz(a), xyz(b));
        // a query ray has no path: its index is its key
        const uint32_t key = alphaRayKey(alphaSalt(0u, 0ull, 0u, 3u), i);
        if (anyHit) {
            uint32_t vis = classifyOccluder(sv, traceAny(sv, r, a.w, b.w, st, key));
            out[i] = make_float4(__uint_as_float(vis), 0.0f, 0.0f, 0.0f);
        } else {
            Hit h = traceClosest(sv, r, a.w, b.w, st, key);
            out[i] = make_float4(h.t, __uint_as_float(h.tri), h.u, h.v);
        }
    }
}

__global__ __launch_bounds__(kTraceBlock) void kPrimary(SceneView sv, SceneData sd, uint32_t rngPolicy,
                                                       uint64_t rngSeed, uint32_t sampleIndex, uint32_t* outMesh,
                                                       uint32_t* outPrim, uint32_t* gstack) {
    __shared__ uint32_t lds[kLdsStackWords];
    TraversalStack st;
    st.lds = lds + threadIdx.x;
    st.glob = gstack + (size_t)(blockIdx.x * kTraceBlock + threadIdx.x) * kGlobalStackWords;
    st.dropped = 0;
    uint32_t vw = sd.viewportRect[2], vh = sd.viewportRect[3];
    uint32_t n = vw * vh;
    for (uint32_t i = blockIdx.x * kTraceBlock + threadIdx.x; i < n; i += gridDim.x * kTraceBlock) {
        int px = (int)(sd.viewportRect[0] + i % vw), py = (int)(sd.viewportRect[1] + i / vw);
        Rng rng = rngInit(rngPolicy, rngSeed, px, py, sd.frameNumber, sampleIndex);
        float jx = rnd(rng);
        float jy = rnd(rng);
        F3 o, d;
        primaryRay(sd, px, py, jx - 0.5f, jy - 0.5f, o, d);
        TraversalRay r = makeTraversalRay(o, d);
        // the key of this pixel's primary ray in the frame's path numbering
        Hit h = traceClosest(sv, r, kRayTMin, kRayTMax, st,
                             alphaRayKey(alphaSalt(sd.frameNumber, rngSeed, 0u, 0u), sampleIndex * n + i));
        if (h.tri == kNoHit) {
            outMesh[i] = VKRT_INVALID_INDEX;
            outPrim[i] = VKRT_INVALID_INDEX;
        } else {
            uint32_t mesh = sv.triMesh[h.tri];
            outMesh[i] = mesh;
            outPrim[i] = h.tri - sv.meshTriBase[mesh];
        }
    }
}

// ------------------------------------------------------------- adjoint --
// Path-replay backpropagation (Vicini, Speierer and Jakob, SIGGRAPH 2021) of
// the RGB integrator. A primal pass stores each path's radiance L; the replay
// regenerates the same path from the same random streams (nothing per vertex
// is stored) and walks L_rem, the radiance still to come, down the path:
//  * every contribution C = throughput * (NEE + emission) the forward pass
//    adds is recomputed and subtracted from L_rem; its own parameter
//    derivatives (light or surface emission, environment texels, the BSDF
//    value at the NEE direction) are accumulated directly;
//  * at a BSDF-sampled continuation the rest of the path scales with the
//    BSDF value f, so d/dtheta adds dI * L_rem * (df/dtheta) / f (sampling and
//    MIS weights are treated as constants, the usual detached estimator).
// Derivatives of f: base colour exactly (the RGB principled BSDF is affine in
// it per channel, so df/db = f(b=1) - f(b=0), times the vertex colour);
// roughness by a central difference of step kRoughnessStep. Shadow rays of the
// replay are traced inline. Gradients are reduced per block in LDS when the
// material table fits, then added to global memory with atomics.
constexpr float kRoughnessStep = 1e-3f;

struct AdjointView {
    const float* dImage;  // accumulation extent, 3 floats per pixel
    float* gBase;         // materialCount * 3
    float* gRough;        // materialCount
    float* gEmission;     // materialCount * 3
    float* gEnv;          // envWidth * envHeight * 3, or null
    const uint32_t* emissiveMeshMaterial;
    uint32_t materialCount;
    uint32_t ldsMaterials;  // material gradients reduced in LDS first
    float invSpp;
};

// slots of one material in the LDS / global gradient layout
struct GradSink {
    float* lds;  // 7 floats per material: base rgb, roughness, emission rgb; null -> global
    const AdjointView* av;
    DEV void add(uint32_t mat, int slot, float v) {
        if (v == 0.0f || mat >= av->materialCount) return;
        if (lds) {
            atomicAdd(&lds[mat * 7u + slot], v);
        } else if (slot < 3) {
            atomicAdd(&av->gBase[mat * 3u + slot], v);
        } else if (slot == 3) {
            atomicAdd(&av->gRough[mat], v);
        } else {
            atomicAdd(&av->gEmission[mat * 3u + (slot - 4)], v);
        }
    }
};

// Texel weights of environmentRadiance's bilinear lookup times the strength:
// adds g (the loss gradient of the returned radiance) to the texels.
DEV void environmentGradient(const SceneView& sv, const SceneData& sd, F3 dirIn, F3 g, float* gEnv) {
    if (!gEnv || sd.environmentTextureIndex == VKRT_INVALID_INDEX || !sv.env) return;
    F3 d = normalize(dirIn);
    float phi = atan2f(d.y, d.x) + sd.environmentRotation * (kPi / 180.0f);
    float theta = acosf(clampf(d.z, -1.0f, 1.0f));
    float u = fracf(phi * (0.5f * kInvPi) + 0.5f);
    float v = theta * kInvPi;
    int W = (int)sv.envWidth, H = (int)sv.envHeight;
    float fx = u * (float)W - 0.5f, fy = v * (float)H - 0.5f;
    float x0f = floorf(fx), y0f = floorf(fy);
    float tx = fx - x0f, ty = fy - y0f;
    int x0 = (int)x0f, y0 = (int)y0f;
    int x1 = x0 + 1, y1 = y0 + 1;
    x0 = wrapOnce(x0, W);
    x1 = wrapOnce(x1, W);
    y0 = min(max(y0, 0), H - 1);
    y1 = min(max(y1, 0), H - 1);
    const float s = sd.environmentLight[3];
    const int xs[4] = {x0, x1, x0, x1}, ys[4] = {y0, y0, y1, y1};
    const float ws[4] = {(1.0f - tx) * (1.0f - ty), tx * (1.0f - ty), (1.0f - tx) * ty, tx * ty};
#pragma unroll
    for (int k = 0; k < 4; k++) {
        float w = ws[k] * s;
        if (w == 0.0f) continue;
        float* t = gEnv + ((size_t)ys[k] * W + xs[k]) * 3u;
        if (g.x != 0.0f) atomicAdd(&t[0], g.x * w);
        if (g.y != 0.0f) atomicAdd(&t[1], g.y * w);
        if (g.z != 0.0f) atomicAdd(&t[2], g.z * w);
    }
}

// BSDF value derivatives at wl: base colour (per channel) and roughness.
struct BsdfDerivatives {
    F3 dBase, dRough;
};
DEV BsdfDerivatives bsdfDerivatives(const Spec& sp, const BMat& mat, F3 vertexColor, F3 wo, uint32_t ff,
                                    float roughnessParam, F3 wl) {
    BsdfDerivatives d;
    BMat m0 = mat, m1 = mat;
    m0.baseColor = f3(0.0f);
    m1.baseColor = f3(1.0f);
    F3 f0 = evalScalar(sp, makeState(m0, wo, ff, 0.0f, 0u), wl).value;
    F3 f1 = evalScalar(sp, makeState(m1, wo, ff, 0.0f, 0u), wl).value;
    d.dBase = (f1 - f0) * vertexColor;
    float rp = sat(roughnessParam + kRoughnessStep), rm = sat(roughnessParam - kRoughnessStep);
    d.dRough = f3(0.0f);
    if (rp > rm) {
        BMat mp = mat, mm = mat;
        mp.roughness = rp;
        mm.roughness = rm;
        F3 fp = evalScalar(sp, makeState(mp, wo, ff, 0.0f, 0u), wl).value;
        F3 fm = evalScalar(sp, makeState(mm, wo, ff, 0.0f, 0u), wl).value;
        d.dRough = (fp - fm) / ((roughnessParam + kRoughnessStep) - (roughnessParam - kRoughnessStep));
    }
    return d;
}

// L_rem of every path starts at the path's primal radiance.
__global__ __launch_bounds__(kBlock) void kAdjointInit(uint32_t count, PathView pv, const float4* sampleRad) {
    uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= count) return;
    pv.rad[i] = sampleRad[i];
}

// One replayed path vertex for every ray of the current queue (RGB mode). The
// forward estimator's operations, random draws and continuation decisions are
// those of kShade<kModeRgb> followed by the shadow passes.
__global__ __launch_bounds__(kTraceBlock) void kShadeAdjoint(FrameParams fp, SceneView sv, PathView pv, QueueView in,
                                                            const float4* hits, QueueView out, uint32_t* counters,
                                                            int inSlot, int outSlot, uint32_t* gstack,
                                                            AdjointView av) {
    __shared__ uint32_t lds[kLdsStackWords];
    extern __shared__ float gradLds[];
    GradSink sink{av.ldsMaterials ? gradLds : nullptr, &av};
    if (av.ldsMaterials) {
        for (uint32_t k = threadIdx.x; k < av.materialCount * 7u; k += kTraceBlock) gradLds[k] = 0.0f;
        __syncthreads();
    }
    TraversalStack st;
    st.lds = lds + threadIdx.x;
    st.glob = gstack + (size_t)(blockIdx.x * kTraceBlock + threadIdx.x) * kGlobalStackWords;
    st.dropped = 0;
    st.sp = 0;
    const uint32_t n = counters[inSlot];
    const uint32_t lane = threadIdx.x & 63u;
    const SceneData& sd = fp.sd;
    for (uint32_t base = blockIdx.x * kTraceBlock + (threadIdx.x & ~63u); base < n; base += gridDim.x * kTraceBlock) {
        const uint32_t i = base + lane;
        bool cont = false;
        F3 nextO = f3(0.0f), nextD = f3(0.0f);
        uint32_t q = 0;
        if (i < n) do {
            float4 ro = in.o[i], rdir = in.d[i];
            q = __float_as_uint(ro.w);
            F3 o = xyz(ro), d = xyz(rdir);
            float4 hh = hits[i];
            Hit h{hh.x, __float_as_uint(hh.y), hh.z, hh.w};
            uint32_t flags = pv.flags[q];
            const bool absorbing = (flags & kFAbsorption) != 0u;
            const bool refractive = (flags & kFRefractive) != 0u;
            F4 thr = ld4(pv.thr[q]);
            F4 misc = ld4(pv.misc[q]);
            F4 lr = ld4(pv.rad[q]);
            F3 lrem = F3{lr.x, lr.y, lr.z};
            // dLoss/d(sample radiance) of this path's pixel
            const uint64_t pg = fp.pathBegin + q;
            const uint32_t k = (uint32_t)(pg % fp.numPixels);
            const uint32_t px = sd.viewportRect[0] + k % fp.vpW, py = sd.viewportRect[1] + k / fp.vpW;
            const float* di = av.dImage + ((size_t)py * fp.accumWidth + px) * 3u;
            const F3 dI = F3{di[0], di[1], di[2]} * av.invSpp;
            const F3 t3 = F3{thr.x, thr.y, thr.z};

            if (h.tri == kNoHit) {
                if (!fp.neeOnly && !(refractive || absorbing)) {
                    F3 env = environmentRadiance(sv, sd, d);
                    const bool prevEnv = (flags & kFPrevEnvNee) != 0u;
                    float envPdf = prevEnv ? environmentPdf(sv, sd, d) : 0.0f;
                    float w = 1.0f;
                    if (prevEnv && envPdf > 0.0f && misc.x > 0.0f) w = powerHeuristic(misc.x, envPdf);
                    environmentGradient(sv, sd, d, dI * t3 * w, av.gEnv);
                }
                break;
            }
            F4 sigma = absorbing ? ld4(pv.sigma[q]) : f4(0.0f);
            if (absorbing) {
                F3 t = transmittance3(sigma, true, h.t);
                thr = F4{thr.x * t.x, thr.y * t.y, thr.z * t.z, thr.w};
            }
            if (!anyPos(F3{thr.x, thr.y, thr.z})) break;
            const F3 thr3 = F3{thr.x, thr.y, thr.z};

            Surface sf = reconstruct(sv, o, d, h);
            const uint32_t mi = sf.materialIndex;
            BMat mat = loadMaterial(sv.materials, mi, sf.vertexColor);
            const float roughParam = sv.materials[mi].roughness;
            F3 wo = toLocal(-d, sf.basis);
            BState bst = makeState(mat, wo, sf.frontFace, 0.0f, 0u);
            const bool curNee = !refractive;
            Rng rng = loadRng(pv.rng[q], fp.rngPolicy);
            bool clearPrevNee = false;

            // light NEE, visibility traced inline
            if (!fp.bsdfOnly && fp.neeEnabled && curNee && wo.z > 0.0f) {
                LightSurfaceSample ls = sampleLightSurface(sv, sd, sf.p, rng);
                if (ls.valid) {
                    F3 off = dot(ls.wi, sf.ng) >= 0.0f ? sf.ng : -sf.ng;
                    F3 wl = toLocal(ls.wi, sf.basis);
                    F3 nee = f3(0.0f);
                    float mis = 0.0f;
                    F3 tr = f3(1.0f);
                    Eval e{};
                    bool evaluated = false;
                    if (!(isRefractive(mat) && wl.z <= 0.0f)) {
                        e = evalScalar(fp.spec, bst, wl);
                        if (e.pdf > 0.0f) {
                            evaluated = true;
                            mis = powerHeuristic(ls.pdfSolidAngle, e.pdf);
                            tr = transmittance3(sigma, absorbing, ls.shadowDistance);
                            F3 fc = e.value * fabsf(wl.z);
                            nee = tr * mis * fc * ls.emission / ls.pdfSolidAngle;
                        }
                    }
                    if (fp.traceShadowAlways || anyPos(nee)) {
                        TraversalRay sr = makeTraversalRay(sf.p + off * kShadowOriginOffset, ls.wi);
                        const uint32_t key =
                            alphaRayKey(alphaSalt(sd.frameNumber, fp.rngSeed, fp.depth, 1u), (uint32_t)fp.pathBegin + q);
                        uint32_t vis = classifyOccluder(sv, traceAny(sv, sr, kRayTMin, ls.shadowDistance, st, key));
                        if (vis == kVisUnsupported) clearPrevNee = true;
                        if (vis == kVisVisible && evaluated) {
                            lrem = lrem - thr3 * nee;
                            // d/d(light emission): thr * tr * mis * f cos / pdf
                            F3 ge = dI * thr3 * tr * (e.value * fabsf(wl.z)) * (mis / ls.pdfSolidAngle);
                            const uint32_t lm = av.emissiveMeshMaterial ? av.emissiveMeshMaterial[ls.light] : 0xFFFFFFFFu;
                            sink.add(lm, 4, ge.x);
                            sink.add(lm, 5, ge.y);
                            sink.add(lm, 6, ge.z);
                            // d/d(surface parameters) through f at the light direction
                            BsdfDerivatives bd = bsdfDerivatives(fp.spec, mat, sf.vertexColor, wo, sf.frontFace,
                                                                 roughParam, wl);
                            F3 g = dI * thr3 * tr * ls.emission * (mis * fabsf(wl.z) / ls.pdfSolidAngle);
                            sink.add(mi, 0, g.x * bd.dBase.x);
                            sink.add(mi, 1, g.y * bd.dBase.y);
                            sink.add(mi, 2, g.z * bd.dBase.z);
                            sink.add(mi, 3, g.x * bd.dRough.x + g.y * bd.dRough.y + g.z * bd.dRough.z);
                        }
                    }
                }
            }

            // environment NEE
            bool envTried = false;
            if (!fp.bsdfOnly && fp.envNee && curNee && !absorbing && wo.z > 0.0f) {
                envTried = true;
                EnvSample es = sampleEnvironmentDirection(sv, sd, rng);
                if (es.valid) {
                    F3 wl = toLocal(es.wi, sf.basis);
                    if (!(isRefractive(mat) && wl.z <= 0.0f)) {
                        F3 le = environmentRadiance(sv, sd, es.wi);
                        Eval e = evalScalar(fp.spec, bst, wl);
                        if (e.pdf > 0.0f) {
                            float mis = powerHeuristic(es.pdf, e.pdf);
                            F3 ne = (e.value * fabsf(wl.z)) * mis * le / es.pdf;
                            if (anyPos(ne)) {
                                F3 off = wl.z > 0.0f ? sf.ng : -sf.ng;
                                TraversalRay sr = makeTraversalRay(sf.p + off * kShadowOriginOffset, es.wi);
                                const uint32_t key = alphaRayKey(alphaSalt(sd.frameNumber, fp.rngSeed, fp.depth, 2u),
                                                                 (uint32_t)fp.pathBegin + q);
                                if (traceAny(sv, sr, kRayTMin, kRayTMax, st, key) == kNoHit) {
                                    lrem = lrem - thr3 * ne;
                                    environmentGradient(sv, sd, es.wi,
                                                        dI * thr3 * (e.value * fabsf(wl.z)) * (mis / es.pdf), av.gEnv);
                                    BsdfDerivatives bd = bsdfDerivatives(fp.spec, mat, sf.vertexColor, wo,
                                                                         sf.frontFace, roughParam, wl);
                                    F3 g = dI * thr3 * le * (mis * fabsf(wl.z) / es.pdf);
                                    sink.add(mi, 0, g.x * bd.dBase.x);
                                    sink.add(mi, 1, g.y * bd.dBase.y);
                                    sink.add(mi, 2, g.z * bd.dBase.z);
                                    sink.add(mi, 3, g.x * bd.dRough.x + g.y * bd.dRough.y + g.z * bd.dRough.z);
                                }
                            }
                        }
                    }
                }
            }

            // emission with MIS. The emission derivative exists on every surface the path
            // hits: a surface that emits nothing is no light source of the current
            // estimator, so only the BSDF strategy reaches it and its weight is 1.
            if (!fp.neeOnly) {
                const bool emissive = anyPos(mat.emission);
                const bool prevNee = (flags & kFPrevNee) != 0u;
                float w = 1.0f;
                if (emissive && fp.neeEnabled && prevNee && fp.depth != 0u && misc.x > 0.0f) {
                    float lp = areaToSolidAngle(sv.meshes[sf.mesh].lightPdfArea, sf.ng, d, h.t);
                    w = lp <= 0.0f ? 1.0f : powerHeuristic(misc.x, lp);
                }
                if (emissive) lrem = lrem - thr3 * (mat.emission * w);
                F3 ge = dI * thr3 * w;
                sink.add(mi, 4, ge.x);
                sink.add(mi, 5, ge.y);
                sink.add(mi, 6, ge.z);
            }

            if (fp.depth + 1u >= sd.rrMaxDepth) break;
            if (wo.z <= 0.0f) break;
            F3 wl;
            uint32_t isT;
            if (!sampleLobe(bst, rng, wl, isT)) break;
            F3 wi = toWorld(wl, sf.basis);
            Eval e = evalScalar(fp.spec, bst, wl);
            if (e.pdf <= 0.0f) break;
            F3 weight = e.value * fabsf(wl.z) / e.pdf;
            if (!(e.pdf > 0.0f && anyPos(weight))) break;
            // the rest of the path (L_rem) scales with f at the sampled direction
            {
                BsdfDerivatives bd = bsdfDerivatives(fp.spec, mat, sf.vertexColor, wo, sf.frontFace, roughParam, wl);
                F3 r = F3{e.value.x > 0.0f ? dI.x * lrem.x / e.value.x : 0.0f,
                          e.value.y > 0.0f ? dI.y * lrem.y / e.value.y : 0.0f,
                          e.value.z > 0.0f ? dI.z * lrem.z / e.value.z : 0.0f};
                sink.add(mi, 0, r.x * bd.dBase.x);
                sink.add(mi, 1, r.y * bd.dBase.y);
                sink.add(mi, 2, r.z * bd.dBase.z);
                sink.add(mi, 3, r.x * bd.dRough.x + r.y * bd.dRough.y + r.z * bd.dRough.z);
            }
            F3 t = thr3 * weight;
            thr = F4{t.x, t.y, t.z, 0.0f};
            if (!anyPos(t)) break;
            misc.x = e.pdf;
            uint32_t newFlags = flags & ~(kFPrevNee | kFPendingCollapse | kFPrevEnvNee);
            if (curNee && isT == 0u) newFlags |= kFPrevNee;
            if (clearPrevNee) newFlags &= ~kFPrevNee;
            if (envTried && isT == 0u) newFlags |= kFPrevEnvNee;
            updateMedium(fp.spec, mat, sf.frontFace, isT, 0.0f, 0u, newFlags, sigma);
            if (fp.depth + 1u >= sd.rrMinDepth) {
                float pr = clampf(maxc(F3{thr.x, thr.y, thr.z}), kRRMin, kRRMax);
                if (rnd(rng) > pr) break;
                F3 t2 = F3{thr.x, thr.y, thr.z} / pr;
                thr = F4{t2.x, t2.y, t2.z, 0.0f};
            }
            F3 offc = isT ? -sf.ng : sf.ng;
            nextO = sf.p + offc * kShadowOriginOffset;
            nextD = wi;
            cont = true;
            pv.rng[q] = storeRng(rng);
            pv.thr[q] = st4(thr);
            pv.misc[q] = st4(misc);
            pv.flags[q] = newFlags;
            if (newFlags & kFAbsorption) pv.sigma[q] = st4(sigma);
            pv.rad[q] = make_float4(lrem.x, lrem.y, lrem.z, 0.0f);
        } while (false);
        uint32_t slot = waveAppend(cont, &counters[outSlot]);
        if (cont) {
            out.o[slot] = make_float4(nextO.x, nextO.y, nextO.z, __uint_as_float(q));
            out.d[slot] = make_float4(nextD.x, nextD.y, nextD.z, kRayTMax);
        }
    }
    if (st.dropped) atomicAdd(&counters[kCntDropped], st.dropped);
    if (av.ldsMaterials) {
        __syncthreads();
        for (uint32_t k = threadIdx.x; k < av.materialCount * 7u; k += kTraceBlock) {
            float v = gradLds[k];
            if (v == 0.0f) continue;
            uint32_t m = k / 7u, slot = k % 7u;
            if (slot < 3u)
                atomicAdd(&av.gBase[m * 3u + slot], v);
            else if (slot == 3u)
                atomicAdd(&av.gRough[m], v);
            else
                atomicAdd(&av.gEmission[m * 3u + (slot - 4u)], v);
        }
    }
}

__global__ __launch_bounds__(kBlock) void kBuildShadeTris(SceneView sv, uint32_t n, ShadeTri* out) {
    uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i < n) out[i] = buildShadeTri(sv, i);
}

// Material class of every triangle for the shading sort: 1 if its material
// transmits (the dielectric lobes), else 0.
__global__ __launch_bounds__(kBlock) void kBuildTriClass(SceneView sv, uint32_t n, uint8_t* cls) {
    uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    const Material& m = sv.materials[sv.meshes[sv.triMesh[i]].materialIndex];
    cls[i] = m.transmission > 0.0f ? 1u : 0u;
}

// Sorting the hits of a depth by material class before shading (waves then
// run one family of BSDF lobes instead of both): class of every queued ray
// (misses count as class 0) and the number of class-0 rays of every block.
__global__ __launch_bounds__(kBlock) void kClassCount(const uint32_t* count, const float4* hits, const uint8_t* triClass,
                                                     uint8_t* key, uint32_t* blockCount0) {
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    const uint32_t n = *count;
    uint32_t c = 1u;
    if (i < n) {
        const uint32_t tri = __float_as_uint(hits[i].y);
        c = tri == kNoHit ? 0u : triClass[tri];
        key[i] = (uint8_t)c;
    }
    const int zeros = __syncthreads_count(i < n && c == 0u);
    if (threadIdx.x == 0) blockCount0[blockIdx.x] = (uint32_t)zeros;
}

// Stable partition: class-0 rays first, then class 1, each in queue order.
// offset0 = exclusive scan of blockCount0; every block but the last is full.
__global__ __launch_bounds__(kBlock) void kClassScatter(const uint32_t* count, const uint8_t* key,
                                                       const uint32_t* blockCount0, const uint32_t* offset0,
                                                       uint32_t blocks, uint32_t* perm) {
    __shared__ uint32_t waveZeros[kBlock / 64];
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    const uint32_t n = *count;
    const uint32_t blockStart = blockIdx.x * kBlock;
    if (blockStart >= n) return;  // uniform per block
    const bool valid = i < n;
    const uint32_t c = valid ? key[i] : 1u;
    const unsigned long long zmask = __ballot(valid && c == 0u);
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    if (lane == 0) waveZeros[wave] = (uint32_t)__popcll(zmask);
    __syncthreads();
    uint32_t before = 0;
    for (uint32_t w = 0; w < wave; w++) before += waveZeros[w];
    const uint32_t rank0 = before + (uint32_t)__popcll(zmask & ((1ull << lane) - 1ull));
    if (!valid) return;
    const uint32_t total0 = offset0[blocks - 1] + blockCount0[blocks - 1];
    uint32_t pos;
    if (c == 0u) {
        pos = offset0[blockIdx.x] + rank0;
    } else {
        const uint32_t rank1 = (i - blockStart) - rank0;  // class-1 rays before i in this block
        pos = total0 + (blockStart - offset0[blockIdx.x]) + rank1;
    }
    perm[pos] = i;
}

// ------------------------------------------------------------- host side --
static uint32_t blocksFor(uint64_t n) { return (uint32_t)((n + kBlock - 1) / kBlock); }

// Per-triangle shading records, rebuilt after a geometry or mesh upload.
bool ensureShadeTris(mrt_context* ctx) {
    if (!ctx->shadeTrisDirty) return true;
    uint32_t n = ctx->triCount;
    if (n > 0) {
        if (!ctx->shadeTris.reserve(n)) return false;
        ctx->shadeTris.count = n;
        hipLaunchKernelGGL(kBuildShadeTris, dim3(blocksFor(n)), dim3(kBlock), 0, ctx->stream, ctx->view(), n,
                           ctx->shadeTris.ptr);
        if (!check(hipGetLastError(), "shading records")) return false;
    }
    ctx->shadeTrisDirty = false;
    return true;
}

// Per-triangle material classes, rebuilt after a geometry, mesh or material upload.
bool ensureTriClass(mrt_context* ctx) {
    if (!ctx->triClassDirty) return true;
    uint32_t n = ctx->triCount;
    if (n > 0) {
        if (!ctx->triClass.reserve(n)) return false;
        hipLaunchKernelGGL(kBuildTriClass, dim3(blocksFor(n)), dim3(kBlock), 0, ctx->stream, ctx->view(), n,
                           ctx->triClass.ptr);
        if (!check(hipGetLastError(), "material classes")) return false;
    }
    ctx->triClassDirty = false;
    return true;
}

bool ensureTraversalScratch(mrt_context* ctx) {
    if (!ctx->counters.reserve(kCounterWords)) return false;
    if (ctx->traceGrid == 0) {
        // persistent traversal grid: enough resident blocks to fill every CU
        // 8 blocks per CU; a multiple of kFetchParts so every claim partition has blocks
        ctx->traceGrid = (uint32_t)std::max(1, ctx->cuCount) * 8u;
        static_assert(8 % kFetchParts == 0, "trace grid must cover every claim partition");
        // the fused tail: two blocks per CU (two waves per SIMD), within the trace grid's stacks
        ctx->tailGrid = (uint32_t)std::max(1, ctx->cuCount) * 2u;
    }
    size_t need = (size_t)ctx->traceGrid * kTraceBlock * kGlobalStackWords;
    return ctx->traversalStack.reserve(need);
}

static bool ensurePaths(mrt_context* ctx, size_t n, int mode, bool lightShadows, bool envShadows) {
    PathBuffers& b = ctx->paths;
    bool ok = b.rng.reserve(n) && b.flags.reserve(n) && b.thr.reserve(n) && b.rad.reserve(n) && b.misc.reserve(n) &&
              b.sigma.reserve(n) && b.hits.reserve(n) && b.sampleRad.reserve(n);
    for (int qi = 0; qi < 2; qi++)
        for (int j = 0; j < 2; j++) ok = ok && b.rays[qi][j].reserve(n);
    if (lightShadows)
        for (int j = 0; j < 4; j++) ok = ok && b.shadow[j].reserve(n);
    if (envShadows)
        for (int j = 0; j < 3; j++) ok = ok && b.shadowEnv[j].reserve(n);
    if (mode != kModeRgb) ok = ok && b.lam.reserve(n);
    if (mode == kModeHero)
        ok = ok && b.radXYZ.reserve(n) && b.tpp.reserve(n) && b.ptpp.reserve(n) && b.pbtp.reserve(n);
    return ok;
}

static PathView pathView(mrt_context* ctx) {
    PathBuffers& b = ctx->paths;
    PathView v;
    v.rng = b.rng.ptr;
    v.flags = b.flags.ptr;
    v.thr = b.thr.ptr;
    v.rad = b.rad.ptr;
    v.radXYZ = b.radXYZ.ptr;
    v.tpp = b.tpp.ptr;
    v.ptpp = b.ptpp.ptr;
    v.pbtp = b.pbtp.ptr;
    v.lam = b.lam.ptr;
    v.misc = b.misc.ptr;
    v.sigma = b.sigma.ptr;
    return v;
}

struct StageEvents {
    std::vector<hipEvent_t> pool;
    std::vector<int> stage;  // per event pair: 0 extend, 1 shade, 2 shadow, 3 cache inference, 4 cache training, 5 fused tail
    std::vector<uint32_t> depth;  // per event pair: the path depth it belongs to
    uint32_t depthNow = 0;
    size_t used = 0;
    bool begin(hipStream_t s, int which) {
        if (used + 2 > pool.size()) return false;
        hipEventRecord(pool[used], s);
        stage.push_back(which);
        depth.push_back(depthNow);
        return true;
    }
    void end(hipStream_t s) {
        hipEventRecord(pool[used + 1], s);
        used += 2;
    }
};

static void queueViews(mrt_context* ctx, QueueView qv[2]) {
    PathBuffers& b = ctx->paths;
    for (int qi = 0; qi < 2; qi++) {
        qv[qi].o = b.rays[qi][0].ptr;
        qv[qi].d = b.rays[qi][1].ptr;
        qv[qi].count = ctx->counters.ptr + qi;
    }
}

// kExtend in the variant the scene needs: counting statistics or not, alpha test or not.
static void launchExtend(mrt_context* ctx, bool countTraversal, const SceneView& sv, const QueueView& q, float4* hits,
                         uint32_t* cnt, int slot, uint32_t salt, uint32_t pathBase,
                         float4 camO = make_float4(0.0f, 0.0f, 0.0f, 0.0f)) {
    hipStream_t s = ctx->stream;
    const bool alpha = sv.triAlpha != nullptr;
    const dim3 grid(ctx->traceGrid), block(kTraceBlock);
    uint32_t* stack = ctx->traversalStack.ptr;
    if (countTraversal && alpha)
        hipLaunchKernelGGL(HIP_KERNEL_NAME(kExtend<true, true, true>), grid, block, 0, s, sv, q.o, q.d, hits, cnt, slot, stack,
                           salt, pathBase, camO);
    else if (countTraversal)
        hipLaunchKernelGGL(HIP_KERNEL_NAME(kExtend<true, false>), grid, block, 0, s, sv, q.o, q.d, hits, cnt, slot, stack,
                           salt, pathBase, camO);
    else if (alpha && sv.texels)
        hipLaunchKernelGGL(HIP_KERNEL_NAME(kExtend<false, true, true>), grid, block, 0, s, sv, q.o, q.d, hits, cnt,
                           slot, stack, salt, pathBase, camO);
    else if (alpha)
        hipLaunchKernelGGL(HIP_KERNEL_NAME(kExtend<false, true, false>), grid, block, 0, s, sv, q.o, q.d, hits, cnt,
                           slot, stack, salt, pathBase, camO);
    else
        hipLaunchKernelGGL(HIP_KERNEL_NAME(kExtend<false, false>), grid, block, 0, s, sv, q.o, q.d, hits, cnt, slot,
                           stack, salt, pathBase, camO);
}

// kShade variant for the scene: RGB with the neural cache, textured, sheen.
template <int MODE, bool SHEEN>
static void launchShade(hipStream_t s, bool tex, const NrcRun* nrc, const FrameParams& fp, const SceneView& sv,
                        const PathView& pv, QueueView in, const float4* hits, QueueView out, ShadowView shv,
                        ShadowView shev, uint32_t* cnt, int cur, int nxt, const uint32_t* order) {
    const dim3 grid(blocksFor(fp.count)), block(kBlock);
    constexpr int NM = MODE == kModeSingle ? kModeRgb : MODE;  // the cache runs in RGB and hero modes
    if (MODE != kModeSingle && nrc && tex)
        hipLaunchKernelGGL(HIP_KERNEL_NAME(kShade<NM, true, true, SHEEN>), grid, block, 0, s, fp, sv, pv, in, hits,
                           out, shv, shev, cnt, cur, nxt, nrc->view, order);
    else if (MODE != kModeSingle && nrc)
        hipLaunchKernelGGL(HIP_KERNEL_NAME(kShade<NM, true, false, SHEEN>), grid, block, 0, s, fp, sv, pv, in, hits,
                           out, shv, shev, cnt, cur, nxt, nrc->view, order);
    else if (tex)
        hipLaunchKernelGGL(HIP_KERNEL_NAME(kShade<MODE, false, true, SHEEN>), grid, block, 0, s, fp, sv, pv, in, hits,
                           out, shv, shev, cnt, cur, nxt, NrcView{}, order);
    else
        hipLaunchKernelGGL(HIP_KERNEL_NAME(kShade<MODE, false, false, SHEEN>), grid, block, 0, s, fp, sv, pv, in,
                           hits, out, shv, shev, cnt, cur, nxt, NrcView{}, order);
}

// tailDepth: the depth from which the fused tail kernel carries the paths left
// (0: none). With traversal statistics the depths stay separate launches and
// only the totals at that depth are recorded.
template <int MODE>
static bool launchChunk(mrt_context* ctx, FrameParams fp, StageEvents& ev, bool countTraversal,
                        const NrcRun* nrc = nullptr, bool keepSampleRad = false, bool sortShading = false,
                        const FeatureView* fv = nullptr, uint32_t tailDepth = 0u) {
    hipStream_t s = ctx->stream;
    PathBuffers& b = ctx->paths;
    PathView pv = pathView(ctx);
    SceneView sv = ctx->view();
    uint32_t* cnt = ctx->counters.ptr;
    QueueView qv[2];
    for (int qi = 0; qi < 2; qi++) {
        qv[qi].o = b.rays[qi][0].ptr;
        qv[qi].d = b.rays[qi][1].ptr;
        qv[qi].count = cnt + qi;
    }
    ShadowView shv{b.shadow[0].ptr, b.shadow[1].ptr, b.shadow[2].ptr, b.shadow[3].ptr, cnt + kCntShadow};
    ShadowView shev{b.shadowEnv[0].ptr, b.shadowEnv[1].ptr, b.shadowEnv[2].ptr, nullptr, cnt + kCntShadowEnv};
    fp.compactPrimary = fp.leanStart && !fv ? 1u : 0u;
    if (fp.leanStart)
        hipLaunchKernelGGL(HIP_KERNEL_NAME(kGenerate<MODE, false>), dim3(blocksFor(fp.count)), dim3(kBlock), 0, s, fp, pv,
                           qv[0]);
    else
        hipLaunchKernelGGL(HIP_KERNEL_NAME(kGenerate<MODE, true>), dim3(blocksFor(fp.count)), dim3(kBlock), 0, s, fp, pv,
                           qv[0]);
    hipLaunchKernelGGL(kSetCount, dim3(1), dim3(1), 0, s, cnt, kCntQueue0, fp.count);
    if (fv) hipLaunchKernelGGL(kFeatInit, dim3(blocksFor(fp.count)), dim3(kBlock), 0, s, fv->state, fp.count);
    int cur = 0;
    uint32_t maxDepth = fp.sd.rrMaxDepth;
    for (uint32_t depth = 0; depth < maxDepth; depth++) {
        int nxt = 1 - cur;
        fp.depth = depth;
        ev.depthNow = depth;
        if (tailDepth != 0u && depth == tailDepth) hipLaunchKernelGGL(kTailSnapshot, dim3(1), dim3(1), 0, s, cnt);
        hipLaunchKernelGGL(kPrepDepth, dim3(1), dim3(64), 0, s, cnt, cur, nxt, depth);
        if (tailDepth != 0u && depth == tailDepth && !countTraversal) {
            bool timed = ev.begin(s, 5);
            hipLaunchKernelGGL(HIP_KERNEL_NAME(kTail<MODE>), dim3(ctx->tailGrid), dim3(kTraceBlock), 0, s, fp, sv, pv,
                               qv[cur], cnt, cur, depth, ctx->traversalStack.ptr);
            if (timed) ev.end(s);
            break;
        }
        bool timed = ev.begin(s, 0);
        const uint32_t pathBase = (uint32_t)fp.pathBegin;
        float4 camO = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        if (fp.compactPrimary && depth == 0u) {
            // primaryRay's origin arithmetic, on the host
            const auto& m = fp.sd.viewInverse;
            float org[3];
            for (int r = 0; r < 3; r++) org[r] = m[0][r] * 0.0f + m[1][r] * 0.0f + m[2][r] * 0.0f + m[3][r] * 1.0f;
            camO = make_float4(org[0], org[1], org[2], 1.0f);
        }
        launchExtend(ctx, countTraversal, sv, qv[cur], b.hits.ptr, cnt, cur,
                     alphaSalt(fp.sd.frameNumber, fp.rngSeed, depth, 0u), pathBase, camO);
        if (timed) ev.end(s);
        timed = ev.begin(s, 1);
        const uint32_t* order = nullptr;
        if (sortShading) {
            const uint32_t nb = blocksFor(fp.count);
            hipLaunchKernelGGL(kClassCount, dim3(nb), dim3(kBlock), 0, s, cnt + cur, b.hits.ptr, ctx->triClass.ptr,
                               b.sortKey.ptr, b.sortCount.ptr);
            size_t bytes = b.sortScratch.capacity;
            if (!check(rocprim::exclusive_scan(b.sortScratch.ptr, bytes, b.sortCount.ptr, b.sortOffset.ptr, 0u, nb,
                                               rocprim::plus<uint32_t>(), s),
                       "shading sort scan"))
                return false;
            hipLaunchKernelGGL(kClassScatter, dim3(nb), dim3(kBlock), 0, s, cnt + cur, b.sortKey.ptr, b.sortCount.ptr,
                               b.sortOffset.ptr, nb, b.sortPerm.ptr);
            order = b.sortPerm.ptr;
        }
        // textured materials take a separate variant so that untextured scenes
        // keep the plain kernel's register budget
        const bool tex = sv.texels != nullptr;
        if (fv) {
            if (tex)
                hipLaunchKernelGGL(HIP_KERNEL_NAME(kFeatures<MODE, true>), dim3(blocksFor(fp.count)), dim3(kBlock), 0,
                                   s, fp, sv, pv, qv[cur], b.hits.ptr, cnt, cur, *fv);
            else
                hipLaunchKernelGGL(HIP_KERNEL_NAME(kFeatures<MODE, false>), dim3(blocksFor(fp.count)), dim3(kBlock),
                                   0, s, fp, sv, pv, qv[cur], b.hits.ptr, cnt, cur, *fv);
        }
        if (ctx->sheenEnabled)
            launchShade<MODE, true>(s, tex, nrc, fp, sv, pv, qv[cur], b.hits.ptr, qv[nxt], shv, shev, cnt, cur, nxt,
                                    order);
        else
            launchShade<MODE, false>(s, tex, nrc, fp, sv, pv, qv[cur], b.hits.ptr, qv[nxt], shv, shev, cnt, cur, nxt,
                                     order);
        if (timed) ev.end(s);
        if (fp.neeEnabled || fp.envNee) {  // otherwise no shadow record can exist
            hipLaunchKernelGGL(kPrepShadow, dim3(1), dim3(1), 0, s, cnt);
            timed = ev.begin(s, 2);
            if (fp.neeEnabled)
                hipLaunchKernelGGL(kShadow<MODE>, dim3(ctx->traceGrid), dim3(kTraceBlock), 0, s, sv, pv, shv, cnt,
                                   kCntShadow, kCntFetchShadow, 0u, ctx->traversalStack.ptr,
                                   alphaSalt(fp.sd.frameNumber, fp.rngSeed, depth, 1u), pathBase);
            if (fp.envNee)
                hipLaunchKernelGGL(kShadow<MODE>, dim3(ctx->traceGrid), dim3(kTraceBlock), 0, s, sv, pv, shev, cnt,
                                   kCntShadowEnv, kCntFetchShadowEnv, 1u, ctx->traversalStack.ptr,
                                   alphaSalt(fp.sd.frameNumber, fp.rngSeed, depth, 2u), pathBase);
            if (timed) ev.end(s);
        }
        cur = nxt;
    }
    uint32_t imagePaths = fp.count;
    if (MODE != kModeSingle && nrc) {
        // the cache's estimate of what image paths would have gathered after their query vertex
        imagePaths = fp.nrcRenderPaths;
        bool timed = ev.begin(s, 3);
        if (!nrcInferQueries(ctx, nrc->view.query, kNrcQueryWords, cnt + kCntNrcQuery, nrc->view.queryCapacity,
                             pv.rad, MODE == kModeHero ? &fp.spec : nullptr))
            return false;
        if (timed) ev.end(s);
        timed = ev.begin(s, 4);
        hipLaunchKernelGGL(kNrcTargets, dim3(blocksFor(std::max(1u, nrc->view.recordCapacity))), dim3(kBlock), 0, s,
                           nrc->view.record, cnt + kCntNrcRecord, nrc->view.recordCapacity, pv.rad, pv.flags,
                           MODE == kModeHero ? 1u : 0u, nrc->view.trainTarget);
        if (nrc->train &&
            !nrcTrainSteps(ctx, nrc->view.record, kNrcRecordWords, nrc->view.trainTarget, cnt + kCntNrcRecord,
                           nrc->view.recordCapacity, nrc->steps, nrc->batch, nrc->learningRate))
            return false;
        if (timed) ev.end(s);
    }
    if (keepSampleRad) {
        hipLaunchKernelGGL(kFinalize<MODE>, dim3(blocksFor(imagePaths)), dim3(kBlock), 0, s, imagePaths, pv,
                           b.sampleRad.ptr);
        hipLaunchKernelGGL(kReduce, dim3(blocksFor(fp.numPixels)), dim3(kBlock), 0, s, fp.numPixels, fp.pathBegin,
                           imagePaths, b.sampleRad.ptr, ctx->frameSum.ptr);
    } else {
        hipLaunchKernelGGL(kFinalizeReduce<MODE>, dim3(blocksFor(fp.numPixels)), dim3(kBlock), 0, s, fp.numPixels,
                           fp.pathBegin, imagePaths, pv, ctx->frameSum.ptr);
    }
    if (fv)
        hipLaunchKernelGGL(kFeatReduce, dim3(blocksFor(fp.numPixels)), dim3(kBlock), 0, s, fp.numPixels, fp.pathBegin,
                           imagePaths, *fv, ctx->frameFeat.ptr);
    return check(hipGetLastError(), "kernel launch");
}

}  // namespace mrt

using namespace mrt;

extern "C" int mrt_resize_accumulation(mrt_context* ctx, uint32_t width, uint32_t height) {
    if (!ctx || width == 0 || height == 0) {
        setError("invalid accumulation extent");
        return -1;
    }
    if (!check(hipSetDevice(ctx->device), "hipSetDevice")) return -2;
    if (ctx->width == width && ctx->height == height && ctx->accum.ptr) return 0;
    size_t n = (size_t)width * height;
    if (!ctx->accum.reserve(n) || !ctx->output.reserve(n)) return -3;
    ctx->accum.count = n;
    ctx->output.count = n;
    ctx->width = width;
    ctx->height = height;
    return mrt_clear_accumulation(ctx);
}

extern "C" int mrt_clear_accumulation(mrt_context* ctx) {
    if (!ctx || !ctx->accum.ptr) return 0;
    size_t n = (size_t)ctx->width * ctx->height;
    hipLaunchKernelGGL(kFill4, dim3(blocksFor(n)), dim3(kBlock), 0, ctx->stream, ctx->accum.ptr, n,
                       make_float4(0.0f, 0.0f, 0.0f, 0.0f));
    // the feature images restart with the accumulation (reallocated on the next feature frame if smaller)
    if (ctx->albedoImg.ptr && ctx->albedoImg.capacity >= n && ctx->normalImg.ptr && ctx->normalImg.capacity >= n) {
        hipLaunchKernelGGL(kFill4, dim3(blocksFor(n)), dim3(kBlock), 0, ctx->stream, ctx->albedoImg.ptr, n,
                           make_float4(0.0f, 0.0f, 0.0f, 0.0f));
        hipLaunchKernelGGL(kFill4, dim3(blocksFor(n)), dim3(kBlock), 0, ctx->stream, ctx->normalImg.ptr, n,
                           make_float4(0.0f, 0.0f, 0.0f, 0.0f));
    }
    ctx->featureImagesValid = false;
    return check(hipGetLastError(), "clear accumulation") ? 0 : -2;
}

extern "C" int mrt_read_features(mrt_context* ctx, float* albedo, float* normal, size_t floatCount) {
    if (!ctx || !albedo || !normal) return -1;
    const size_t n = (size_t)ctx->width * ctx->height;
    if (!ctx->featureImagesValid || floatCount < n * 4u) {
        setError("no feature images (enable them before rendering) or buffer too small");
        return -1;
    }
    if (!check(hipSetDevice(ctx->device), "hipSetDevice") || !check(hipStreamSynchronize(ctx->stream), "sync") ||
        !check(hipMemcpy(albedo, ctx->albedoImg.ptr, n * sizeof(float4), hipMemcpyDeviceToHost), "read albedo") ||
        !check(hipMemcpy(normal, ctx->normalImg.ptr, n * sizeof(float4), hipMemcpyDeviceToHost), "read normal"))
        return -2;
    return 0;
}

// Validation, buffers and per-frame parameters shared by the forward and the
// adjoint passes.
struct FrameSetup {
    FrameParams fp;
    uint64_t totalPaths = 0, chunk = 0;
    int mode = kModeRgb;
    uint32_t renderMode = 0, vw = 0, vh = 0;
    uint32_t nrcQueryCapacity = 0, nrcRecordCapacity = 0;
    bool sortShading = false;
};

static int setupFrame(mrt_context* ctx, const SceneData* scene, const mrt_render_params* prm, FrameSetup& fs) {
    if (!check(hipSetDevice(ctx->device), "hipSetDevice")) return -2;
    if (mrt_resize_accumulation(ctx, prm->width, prm->height) != 0) return -2;
    if (!ctx->rgb2spec.ptr) {
        setError("rgb2spec table not uploaded");
        return -2;
    }
    if (scene->rgb2specSRGB.res != kSpecRes) {
        setError("the shading kernels are built for a 64^3 rgb2spec table");
        return -2;
    }
    if (ctx->triCount > 0 && ctx->bvhMaxDepth * (kBvhWidth - 1u) + 1u > (uint32_t)(kLdsStack + kGlobalStack)) {
        setError("BVH deeper than the traversal stack supports");
        return -2;
    }
    hipStream_t s = ctx->stream;
    uint32_t vx = scene->viewportRect[0], vy = scene->viewportRect[1];
    uint32_t vw = scene->viewportRect[2], vh = scene->viewportRect[3];
    if (vx + vw > prm->width) vw = prm->width > vx ? prm->width - vx : 0;
    if (vy + vh > prm->height) vh = prm->height > vy ? prm->height - vy : 0;
    uint32_t renderMode = VKRT_SCENE_RENDER_MODE_FROM_RENDER_SETTINGS(scene->packedRenderSettings);
    uint32_t specMode = VKRT_SCENE_SPECTRAL_SAMPLING_MODE_FROM_RENDER_SETTINGS(scene->packedRenderSettings);
    int mode = renderMode == VKRT_RENDER_MODE_SPECTRAL ? (specMode == VKRT_SPECTRAL_SAMPLING_MODE_HERO ? kModeHero
                                                                                                      : kModeSingle)
                                                       : kModeRgb;
    uint32_t spp = scene->samplesPerPixel > 1u ? scene->samplesPerPixel : 1u;
    uint64_t numPixels = (uint64_t)vw * vh;
    if (prm->resetAccumulation && mrt_clear_accumulation(ctx) != 0) return -2;
    if (!ensureTraversalScratch(ctx) || !ensureShadeTris(ctx)) return -3;
    if (!ctx->frameSum.reserve(numPixels ? numPixels : 1)) return -3;
    hipLaunchKernelGGL(kFill4, dim3(blocksFor(numPixels ? numPixels : 1)), dim3(kBlock), 0, s, ctx->frameSum.ptr,
                       (size_t)numPixels, make_float4(0.0f, 0.0f, 0.0f, 0.0f));
    hipMemsetAsync(ctx->counters.ptr, 0, kCounterWords * sizeof(uint32_t), s);

    uint64_t totalPaths = numPixels * spp;
    uint64_t chunk = std::min<uint64_t>(totalPaths, kMaxChunk);
    if (numPixels > 0 && numPixels <= kMaxChunk) chunk = std::min<uint64_t>(totalPaths, (kMaxChunk / numPixels) * numPixels);
    const bool nrc = prm->nrcEnabled != 0u && (mode == kModeRgb || mode == kModeHero) && numPixels > 0;
    if (nrc) {
        // one chunk: the image paths followed by the cache's training paths
        if (totalPaths + prm->nrcTrainingPaths > kMaxChunk) {
            setError("neural cache frames must fit one chunk of paths");
            return -2;
        }
        chunk = totalPaths + prm->nrcTrainingPaths;
        if (!nrcEnsure(ctx, 0)) return -3;
    }
    fs.sortShading = prm->materialSort != 0u && ctx->triCount > 0;
    if (fs.sortShading && chunk > 0) {
        PathBuffers& b = ctx->paths;
        const uint32_t nb = blocksFor(chunk);
        size_t scratch = 0;
        if (!ensureTriClass(ctx) || !b.sortKey.reserve(chunk) || !b.sortPerm.reserve(chunk) ||
            !b.sortCount.reserve(nb) || !b.sortOffset.reserve(nb) ||
            !check(rocprim::exclusive_scan(nullptr, scratch, b.sortCount.ptr, b.sortOffset.ptr, 0u, nb,
                                           rocprim::plus<uint32_t>(), ctx->stream),
                   "shading sort scan size") ||
            !b.sortScratch.reserve(scratch))
            return -3;
    }
    FrameParams& fp = fs.fp;
    memset(&fp, 0, sizeof(fp));
    if (nrc) {
        fp.nrc = 1u;
        fp.nrcPrefix = prm->nrcPrefix;
        fp.nrcRenderPaths = (uint32_t)totalPaths;
        for (int a = 0; a < 3; a++) {
            float ext = prm->sceneMax[a] - prm->sceneMin[a];
            fp.nrcBoxMin[a] = prm->sceneMin[a];
            fp.nrcBoxInv[a] = ext > 0.0f ? 1.0f / ext : 0.0f;
        }
        const size_t records = (size_t)prm->nrcTrainingPaths * std::max(1u, scene->rrMaxDepth);
        PathBuffers& b = ctx->paths;
        if (!b.nrcQuery.reserve(totalPaths * kNrcQueryWords) || !b.nrcRecord.reserve(std::max<size_t>(records, 1) *
                                                                                      kNrcRecordWords) ||
            !b.nrcTarget.reserve(std::max<size_t>(records, 1)))
            return -3;
        fs.nrcQueryCapacity = (uint32_t)totalPaths;
        fs.nrcRecordCapacity = (uint32_t)records;
    }
    fp.sd = *scene;
    fp.numPixels = (uint32_t)numPixels;
    fp.vpW = vw;
    fp.accumWidth = prm->width;
    fp.rngPolicy = prm->rngPolicy;
    fp.rngSeed = prm->rngSeed;
    fp.neeEnabled = scene->misNeeEnabled != 0u && scene->emissiveMeshCount > 0u;
    fp.bsdfOnly = scene->debugMode == VKRT_DEBUG_MODE_BSDF_ONLY;
    fp.neeOnly = scene->debugMode == VKRT_DEBUG_MODE_NEE_ONLY;
    fp.traceShadowAlways = prm->hasTransmissive ? 1u : 0u;
    fp.envNee = prm->envImportance != 0u && ctx->envWidth > 0u && ctx->envAliasQ.ptr &&
                scene->environmentTextureIndex != VKRT_INVALID_INDEX;
    if (chunk > 0 && !ensurePaths(ctx, chunk, mode, fp.neeEnabled != 0u, fp.envNee != 0u)) return -3;
    fp.leanStart = 1u;
    fp.accum = ctx->accum.ptr;
    fp.spec.table = ctx->rgb2spec.ptr;
    fp.spec.res = scene->rgb2specSRGB.res;
    fp.spec.scaleOffset = scene->rgb2specSRGB.scaleOffset;
    fp.spec.dataOffset = scene->rgb2specSRGB.dataOffset;
    fs.totalPaths = totalPaths;
    fs.chunk = chunk;
    fs.mode = mode;
    fs.renderMode = renderMode;
    fs.vw = vw;
    fs.vh = vh;
    return 0;
}

extern "C" int mrt_render_frame(mrt_context* ctx, const SceneData* scene, const mrt_render_params* prm,
                                mrt_render_stats* stats) {
    if (!ctx || !scene || !prm) {
        setError("mrt_render_frame: null argument");
        return -1;
    }
    auto wall0 = std::chrono::steady_clock::now();
    FrameSetup fs;
    int rc = setupFrame(ctx, scene, prm, fs);
    if (rc != 0) return rc;
    hipStream_t s = ctx->stream;
    FrameParams& fp = fs.fp;
    const uint64_t totalPaths = fs.totalPaths, chunk = fs.chunk;
    const int mode = fs.mode;
    const uint32_t renderMode = fs.renderMode, vw = fs.vw, vh = fs.vh;

    static StageEvents ev;
    if (ev.pool.empty()) {
        ev.pool.resize(8192);
        for (auto& e : ev.pool) hipEventCreate(&e);
    }
    ev.used = 0;
    ev.stage.clear();
    const bool countTraversal = prm->traversalStatistics != 0u;
    hipEventRecord(ctx->evBegin, s);
    uint32_t chunks = 0;
    NrcRun nrcRun{};
    if (fp.nrc) {
        PathBuffers& b = ctx->paths;
        nrcRun.view = NrcView{b.nrcQuery.ptr, b.nrcRecord.ptr, nullptr, b.nrcTarget.ptr, fs.nrcQueryCapacity,
                              fs.nrcRecordCapacity};
        nrcRun.train = prm->nrcTrain != 0u;
        nrcRun.steps = prm->nrcTrainingSteps;
        nrcRun.batch = prm->nrcBatch;
        nrcRun.learningRate = prm->nrcLearningRate;
    }
    // feature images and debug views: a feature pass per depth, then the
    // accumulation kernel that writes them (off for the plain render)
    const uint32_t dm = scene->debugMode;
    const bool debugView = dm != VKRT_DEBUG_MODE_NONE && dm != VKRT_DEBUG_MODE_NEE_ONLY &&
                           dm != VKRT_DEBUG_MODE_BSDF_ONLY;
    const bool featurePass = prm->features != 0u || (debugView && dm != VKRT_DEBUG_MODE_SELECTION_MASK);
    const size_t n = (size_t)prm->width * prm->height;
    FeatureView fview{};
    if (featurePass || debugView) {
        PathBuffers& b = ctx->paths;
        const size_t px = std::max<size_t>(fp.numPixels, 1);
        if (!b.featState.reserve(chunk) || !b.featAlbedo.reserve(chunk) || !b.featNormal.reserve(chunk) ||
            !ctx->frameFeat.reserve(3 * px) || !ctx->debug0.reserve(px) || !ctx->bounce0.reserve(px))
            return -3;
        if (!ctx->albedoImg.ptr || ctx->albedoImg.capacity < n || !ctx->normalImg.ptr || ctx->normalImg.capacity < n) {
            if (!ctx->albedoImg.reserve(n) || !ctx->normalImg.reserve(n)) return -3;
            hipLaunchKernelGGL(kFill4, dim3(blocksFor(n)), dim3(kBlock), 0, s, ctx->albedoImg.ptr, n,
                               make_float4(0.0f, 0.0f, 0.0f, 0.0f));
            hipLaunchKernelGGL(kFill4, dim3(blocksFor(n)), dim3(kBlock), 0, s, ctx->normalImg.ptr, n,
                               make_float4(0.0f, 0.0f, 0.0f, 0.0f));
        }
        hipLaunchKernelGGL(kFill4, dim3(blocksFor(3 * px)), dim3(kBlock), 0, s, ctx->frameFeat.ptr, 3 * px,
                           make_float4(0.0f, 0.0f, 0.0f, 0.0f));
        hipLaunchKernelGGL(kFill4, dim3(blocksFor(px)), dim3(kBlock), 0, s, ctx->debug0.ptr, px,
                           make_float4(0.0f, 0.0f, 0.0f, 0.0f));
        hipLaunchKernelGGL(kFeatInit, dim3(blocksFor(px)), dim3(kBlock), 0, s, ctx->bounce0.ptr, (uint32_t)px);
        fview = FeatureView{b.featState.ptr, b.featAlbedo.ptr, b.featNormal.ptr, ctx->debug0.ptr, ctx->bounce0.ptr};
    }
    const FeatureView* fv = featurePass ? &fview : nullptr;
    // The fused tail (kTail) takes over from the first depth whose queue held
    // at most prm->tailRays rays in the previous frame of the same size: a
    // single-chunk frame whose shading variant is the plain one (no shadow
    // rays, cache, features, material sort, alpha test, textures or sheen).
    uint32_t tailDepth = 0u;
    const uint32_t maxDepth = scene->rrMaxDepth;
    const SceneView svNow = ctx->view();
    const bool tailAllowed = prm->tailRays > 0u && totalPaths == chunk && !fp.nrc && !fv && !fs.sortShading &&
                             !fp.neeEnabled && !fp.envNee && svNow.triAlpha == nullptr && svNow.texels == nullptr &&
                             !ctx->sheenEnabled;
    if (tailAllowed && ctx->hintPaths == chunk && ctx->hintMaxDepth == maxDepth) {
        for (uint32_t d = 1; d < std::min<uint32_t>(maxDepth, kDepthHist); d++)
            if (ctx->depthHint[d] <= prm->tailRays) {
                tailDepth = d;
                break;
            }
    }
    for (uint64_t begin = 0; begin < totalPaths; begin += chunk) {
        fp.pathBegin = begin;
        fp.count = (uint32_t)std::min<uint64_t>(chunk, totalPaths - begin);
        if (fp.nrc) fp.count = (uint32_t)chunk;  // image paths and training paths in one chunk
        const bool sorted = fs.sortShading;
        bool ok = mode == kModeRgb
                      ? launchChunk<kModeRgb>(ctx, fp, ev, countTraversal, fp.nrc ? &nrcRun : nullptr, false, sorted,
                                              fv, tailDepth)
                  : mode == kModeHero ? launchChunk<kModeHero>(ctx, fp, ev, countTraversal, fp.nrc ? &nrcRun : nullptr,
                                                               false, sorted, fv, tailDepth)
                                      : launchChunk<kModeSingle>(ctx, fp, ev, countTraversal, nullptr, false, sorted,
                                                                 fv, tailDepth);
        if (!ok) return -4;
        chunks++;
    }
    const uint32_t spectralOut = renderMode == VKRT_RENDER_MODE_SPECTRAL ? 1u : 0u;
    if (featurePass || debugView) {
        hipLaunchKernelGGL(kAccumulateFeatures, dim3(blocksFor(n)), dim3(kBlock), 0, s, fp.sd, prm->width, prm->height,
                           vw, vh, ctx->frameSum.ptr, ctx->frameFeat.ptr, fview, ctx->accum.ptr, ctx->albedoImg.ptr,
                           ctx->normalImg.ptr, ctx->output.ptr, spectralOut, featurePass ? 1u : 0u);
        ctx->featureImagesValid = true;
    } else {
        hipLaunchKernelGGL(kAccumulate, dim3(blocksFor(n)), dim3(kBlock), 0, s, fp.sd, prm->width, prm->height, vw,
                           vh, ctx->frameSum.ptr, ctx->accum.ptr, ctx->output.ptr, spectralOut);
    }
    hipEventRecord(ctx->evEnd, s);
    if (!check(hipEventSynchronize(ctx->evEnd), "frame")) return -4;
    uint32_t hostCounters[64];
    if (!check(hipMemcpy(hostCounters, ctx->counters.ptr, sizeof(hostCounters), hipMemcpyDeviceToHost), "counters"))
        return -4;
    if (hostCounters[kCntDropped]) {
        setError("traversal stack overflow");
        return -5;
    }
    if (totalPaths == chunk && !fp.nrc) {
        // queue lengths for the next frame's choice of the tail depth (depths
        // the fused tail carried were not recorded and read as 0)
        for (uint32_t d = 0; d < (uint32_t)kDepthHist; d++) ctx->depthHint[d] = hostCounters[kCntDepthHist + d];
        ctx->hintPaths = chunk;
        ctx->hintMaxDepth = maxDepth;
    }
    if (stats) {
        unsigned long long st64[8];
        memcpy(st64, hostCounters + kCntStats, sizeof(st64));
        stats->extensionRays = st64[0];
        stats->shadowRays = st64[1];
        stats->nodeVisits = st64[2];
        stats->triangleTests = st64[3];
        stats->paths = totalPaths;
        float ms = 0.0f;
        hipEventElapsedTime(&ms, ctx->evBegin, ctx->evEnd);
        stats->traceMs = ms;
        float stageMs[6] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
        uint32_t stageN[6] = {0u, 0u, 0u, 0u, 0u, 0u};
        for (size_t e = 0; e + 1 < ev.used; e += 2) {
            float t = 0.0f;
            hipEventElapsedTime(&t, ev.pool[e], ev.pool[e + 1]);
            int which = ev.stage[e / 2];
            stageMs[which] += t;
            stageN[which]++;
            if (ev.depth[e / 2] == 0u && which == 0) stats->extendMsDepth0 += t;
            if (ev.depth[e / 2] == 0u && which == 1) stats->shadeMsDepth0 += t;
        }
        stats->extendMs = stageMs[0];
        stats->shadeMs = stageMs[1];
        stats->shadowMs = stageMs[2];
        stats->extendLaunches = stageN[0];
        stats->shadeLaunches = stageN[1];
        stats->shadowLaunches = stageN[2];
        stats->nrcInferMs = stageMs[3];
        stats->nrcTrainMs = stageMs[4];
        stats->raysBeforeTail = tailDepth ? st64[5] : st64[0];
        stats->nodeVisitsBeforeTail = tailDepth ? st64[6] : st64[2];
        stats->triangleTestsBeforeTail = tailDepth ? st64[7] : st64[3];
        if (tailDepth && !countTraversal) {
            stats->tailStartDepth = tailDepth;
            stats->tailLaunches = stageN[5];
            stats->tailMs = stageMs[5];
            stats->tailRays = st64[0] - st64[5];
        }
        if (fp.nrc) {
            stats->nrcQueries = std::min(hostCounters[kCntNrcQuery], fs.nrcQueryCapacity);
            stats->nrcRecords = std::min(hostCounters[kCntNrcRecord], fs.nrcRecordCapacity);
            const float denom = (float)prm->nrcTrainingSteps * (float)prm->nrcBatch * 3.0f;
            stats->nrcLoss = prm->nrcTrain && denom > 0.0f ? nrcReadLoss(ctx) / denom : 0.0f;
        }
        (void)wall0;
    }
    return 0;
}

extern "C" int mrt_render_backward(mrt_context* ctx, const SceneData* scene, const mrt_render_params* prm,
                                   const float* dImage, mrt_gradients* g, mrt_render_stats* stats) {
    if (!ctx || !scene || !prm || !dImage || !g) {
        setError("mrt_render_backward: null argument");
        return -1;
    }
    mrt_render_params p2 = *prm;
    p2.resetAccumulation = 0;
    FrameSetup fs;
    int rc = setupFrame(ctx, scene, &p2, fs);
    if (rc != 0) return rc;
    if (fs.mode != kModeRgb) {
        setError("render_backward supports the RGB render mode");
        return -1;
    }
    if (ctx->texturesEnabled) {
        setError("render_backward does not differentiate through material textures");
        return -1;
    }
    if (g->materialCount != (uint32_t)ctx->materials.count) {
        setError("render_backward: gradient material count differs from the scene");
        return -1;
    }
    hipStream_t s = ctx->stream;
    FrameParams& fp = fs.fp;
    const uint32_t M = g->materialCount;
    const bool envGrad = g->environment && ctx->envWidth > 0u && g->environmentWidth == ctx->envWidth &&
                         g->environmentHeight == ctx->envHeight &&
                         scene->environmentTextureIndex != VKRT_INVALID_INDEX;
    const size_t envN = envGrad ? (size_t)ctx->envWidth * ctx->envHeight * 3u : 0u;
    const size_t pix = (size_t)prm->width * prm->height;
    if (!ctx->adjImage.upload(dImage, pix * 3u, s) || !ctx->adjGrad.reserve((size_t)M * 7u + envN + 1u)) return -3;
    hipMemsetAsync(ctx->adjGrad.ptr, 0, ((size_t)M * 7u + envN) * sizeof(float), s);
    AdjointView av;
    av.dImage = ctx->adjImage.ptr;
    av.gBase = ctx->adjGrad.ptr;
    av.gRough = ctx->adjGrad.ptr + (size_t)M * 3u;
    av.gEmission = ctx->adjGrad.ptr + (size_t)M * 4u;
    av.gEnv = envGrad ? ctx->adjGrad.ptr + (size_t)M * 7u : nullptr;
    av.emissiveMeshMaterial = scene->emissiveMeshCount ? ctx->emissiveMeshMaterial.ptr : nullptr;
    av.materialCount = M;
    const size_t gradLdsBytes = (size_t)M * 7u * sizeof(float);
    av.ldsMaterials = gradLdsBytes <= 16384u ? 1u : 0u;
    uint32_t spp = scene->samplesPerPixel > 1u ? scene->samplesPerPixel : 1u;
    av.invSpp = 1.0f / (float)spp;

    PathBuffers& b = ctx->paths;
    PathView pv = pathView(ctx);
    SceneView sv = ctx->view();
    uint32_t* cnt = ctx->counters.ptr;
    QueueView qv[2];
    queueViews(ctx, qv);
    StageEvents untimed;
    hipEventRecord(ctx->evBegin, s);
    for (uint64_t begin = 0; begin < fs.totalPaths; begin += fs.chunk) {
        fp.pathBegin = begin;
        fp.count = (uint32_t)std::min<uint64_t>(fs.chunk, fs.totalPaths - begin);
        // primal: the radiance of every path of the chunk (sampleRad)
        if (!launchChunk<kModeRgb>(ctx, fp, untimed, false, nullptr, true)) return -4;
        // replay with the same streams
        hipLaunchKernelGGL(HIP_KERNEL_NAME(kGenerate<kModeRgb, true>), dim3(blocksFor(fp.count)), dim3(kBlock), 0, s, fp,
                           pv, qv[0]);
        hipLaunchKernelGGL(kSetCount, dim3(1), dim3(1), 0, s, cnt, kCntQueue0, fp.count);
        hipLaunchKernelGGL(kAdjointInit, dim3(blocksFor(fp.count)), dim3(kBlock), 0, s, fp.count, pv,
                           b.sampleRad.ptr);
        int cur = 0;
        for (uint32_t depth = 0; depth < fp.sd.rrMaxDepth; depth++) {
            int nxt = 1 - cur;
            fp.depth = depth;
            hipLaunchKernelGGL(kPrepDepth, dim3(1), dim3(64), 0, s, cnt, cur, nxt, depth);
            // the same alpha decisions as the primal pass: same salt, same path numbering
            launchExtend(ctx, false, sv, qv[cur], b.hits.ptr, cnt, cur, alphaSalt(fp.sd.frameNumber, fp.rngSeed, depth, 0u),
                         (uint32_t)fp.pathBegin);
            hipLaunchKernelGGL(kShadeAdjoint, dim3(ctx->traceGrid), dim3(kTraceBlock),
                               av.ldsMaterials ? (uint32_t)gradLdsBytes : 0u, s, fp, sv, pv, qv[cur], b.hits.ptr,
                               qv[nxt], cnt, cur, nxt, ctx->traversalStack.ptr, av);
            cur = nxt;
        }
        if (!check(hipGetLastError(), "adjoint launch")) return -4;
    }
    hipEventRecord(ctx->evEnd, s);
    std::vector<float> host((size_t)M * 7u + envN);
    if (!host.empty() &&
        !check(hipMemcpyAsync(host.data(), ctx->adjGrad.ptr, host.size() * sizeof(float), hipMemcpyDeviceToHost, s),
               "gradients"))
        return -4;
    uint32_t hostCounters[64];
    if (!check(hipMemcpyAsync(hostCounters, cnt, sizeof(hostCounters), hipMemcpyDeviceToHost, s), "counters") ||
        !check(hipStreamSynchronize(s), "render backward"))
        return -4;
    if (hostCounters[kCntDropped]) {
        setError("traversal stack overflow");
        return -5;
    }
    if (g->baseColor) memcpy(g->baseColor, host.data(), (size_t)M * 3u * sizeof(float));
    if (g->roughness) memcpy(g->roughness, host.data() + (size_t)M * 3u, (size_t)M * sizeof(float));
    if (g->emission) memcpy(g->emission, host.data() + (size_t)M * 4u, (size_t)M * 3u * sizeof(float));
    if (envGrad) memcpy(g->environment, host.data() + (size_t)M * 7u, envN * sizeof(float));
    if (stats) {
        memset(stats, 0, sizeof(*stats));
        unsigned long long st64[2];
        memcpy(st64, hostCounters + kCntStats, sizeof(st64));
        stats->extensionRays = st64[0];
        stats->shadowRays = st64[1];
        stats->paths = fs.totalPaths;
        float ms = 0.0f;
        hipEventElapsedTime(&ms, ctx->evBegin, ctx->evEnd);
        stats->traceMs = ms;
    }
    return 0;
}

extern "C" int mrt_read_accumulation(mrt_context* ctx, float* out, size_t floatCount) {
    if (!ctx || !out || floatCount != (size_t)ctx->width * ctx->height * 4u) {
        setError("mrt_read_accumulation: size mismatch");
        return -1;
    }
    hipStreamSynchronize(ctx->stream);
    return check(hipMemcpy(out, ctx->accum.ptr, floatCount * sizeof(float), hipMemcpyDeviceToHost), "read accum") ? 0 : -2;
}

extern "C" int mrt_read_output(mrt_context* ctx, uint16_t* out, size_t count) {
    if (!ctx || !out || count != (size_t)ctx->width * ctx->height * 4u) {
        setError("mrt_read_output: size mismatch");
        return -1;
    }
    hipStreamSynchronize(ctx->stream);
    return check(hipMemcpy(out, ctx->output.ptr, count * sizeof(uint16_t), hipMemcpyDeviceToHost), "read output") ? 0 : -2;
}

static uint32_t hostBits(float f) {
    uint32_t u;
    memcpy(&u, &f, sizeof(u));
    return u;
}

static int runQuery(mrt_context* ctx, const float* rays, size_t count, uint32_t anyHit, std::vector<float4>& result) {
    if (!check(hipSetDevice(ctx->device), "hipSetDevice")) return -2;
    if (!ensureTraversalScratch(ctx)) return -3;
    if (count > 0xFFFFFFFFull / 2) {
        setError("too many rays");
        return -1;
    }
    float4* dRays = nullptr;
    float4* dOut = nullptr;
    if (!check(hipMalloc(&dRays, count * 2 * sizeof(float4)), "hipMalloc rays")) return -3;
    if (!check(hipMalloc(&dOut, count * sizeof(float4)), "hipMalloc out")) {
        hipFree(dRays);
        return -3;
    }
    int rc = 0;
    if (check(hipMemcpy(dRays, rays, count * 8 * sizeof(float), hipMemcpyHostToDevice), "upload rays")) {
        hipLaunchKernelGGL(kTraceQuery, dim3(ctx->traceGrid), dim3(kTraceBlock), 0, ctx->stream, ctx->view(), dRays,
                           (uint32_t)count, dOut, ctx->traversalStack.ptr, anyHit);
        result.resize(count);
        if (!check(hipStreamSynchronize(ctx->stream), "trace query") ||
            !check(hipMemcpy(result.data(), dOut, count * sizeof(float4), hipMemcpyDeviceToHost), "read hits"))
            rc = -4;
    } else {
        rc = -4;
    }
    hipFree(dRays);
    hipFree(dOut);
    return rc;
}

// ------------------------------------------------------ allocation-free --
// The allocation-free ray query keeps everything beyond its inputs and
// outputs in one fixed device array: a header whose word 0 counts dropped
// traversal-stack pushes, then the global traversal stacks of a persistent
// grid of kNoAllocBlocks blocks (kTraceBlock lanes x kGlobalStackWords words
// each): 64 + 256 * 256 * 232 words, about 61 MB, held on every device the
// library runs on until the process ends.
constexpr uint32_t kNoAllocBlocks = 256;
constexpr size_t kNoAllocHeaderWords = 64;
constexpr size_t kNoAllocScratchWords =
    kNoAllocHeaderWords + (size_t)kNoAllocBlocks * kTraceBlock * (size_t)kGlobalStackWords;
static __device__ __align__(256) uint32_t scratch[kNoAllocScratchWords];

// Device address of scratch on the current device; called once per context
// at creation with its device current (the array's name is not a device
// address in host code).
bool mrt::noAllocScratchInit(mrt_context* ctx) {
    void* p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) {
        (void)hipGetLastError();
        ctx->noAllocScratch = nullptr;
        return false;
    }
    ctx->noAllocScratch = static_cast<uint32_t*>(p);
    return true;
}

__global__ void kClearWords(uint32_t* p, uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0u;
}

// kTraceQuery's closest-hit search (same keys, same stack) writing the
// per-mesh answer directly: mesh, primitive within the mesh, t, barycentrics.
__global__ __launch_bounds__(kTraceBlock) void kTraceQueryDirect(SceneView sv, const float4* rays, uint32_t n,
                                                                const uint32_t* triMesh, const uint32_t* meshTriBase,
                                                                uint32_t* outMesh, uint32_t* outPrim, float* outT,
                                                                float* outBary, uint32_t* scratchWords) {
    __shared__ uint32_t lds[kLdsStackWords];
    TraversalStack st;
    st.lds = lds + threadIdx.x;
    st.glob = scratchWords + kNoAllocHeaderWords + (size_t)(blockIdx.x * kTraceBlock + threadIdx.x) * kGlobalStackWords;
    st.dropped = 0;
    for (uint32_t i = blockIdx.x * kTraceBlock + threadIdx.x; i < n; i += gridDim.x * kTraceBlock) {
        const float4 a = rays[2 * (size_t)i], b = rays[2 * (size_t)i + 1];
        const TraversalRay r = makeTraversalRay(xyz(a), xyz(b));
        const Hit h = traceClosest(sv, r, a.w, b.w, st, alphaRayKey(alphaSalt(0u, 0ull, 0u, 3u), i));
        const bool hit = h.tri != kNoHit && h.tri < sv.triCount;
        const uint32_t mesh = hit ? triMesh[h.tri] : VKRT_INVALID_INDEX;
        outMesh[i] = mesh;
        outPrim[i] = hit ? h.tri - meshTriBase[mesh] : VKRT_INVALID_INDEX;
        if (outT) outT[i] = hit ? h.t : 0.0f;
        if (outBary) {
            outBary[2 * (size_t)i] = hit ? h.u : 0.0f;
            outBary[2 * (size_t)i + 1] = hit ? h.v : 0.0f;
        }
    }
    if (st.dropped) atomicAdd(&scratchWords[0], st.dropped);
}

extern "C" int mrt_trace_rays_noalloc(mrt_context* ctx, const float* deviceRays, size_t count, uint32_t* deviceMesh,
                                      uint32_t* devicePrim, float* deviceT, float* deviceBary) {
    if (!ctx || !deviceRays || !deviceMesh || !devicePrim) return -1;
    if (count > MRT_NOALLOC_MAX_RAYS) {
        setError("allocation-free ray query: more rays than its maximum of 2^30 per call");
        return -1;
    }
    if (!ctx->noAllocScratch) {
        setError("allocation-free ray query: no device address for its scratch array");
        return -2;
    }
    if (!check(hipSetDevice(ctx->device), "hipSetDevice")) return -2;
    if (count == 0) return 0;
    hipStream_t s = ctx->stream;
    hipLaunchKernelGGL(kClearWords, dim3(1), dim3((uint32_t)kNoAllocHeaderWords), 0, s, ctx->noAllocScratch,
                       (uint32_t)kNoAllocHeaderWords);
    hipLaunchKernelGGL(kTraceQueryDirect, dim3(kNoAllocBlocks), dim3(kTraceBlock), 0, s, ctx->view(),
                       reinterpret_cast<const float4*>(deviceRays), (uint32_t)count, ctx->triMesh.ptr,
                       ctx->meshTriBase.ptr, deviceMesh, devicePrim, deviceT, deviceBary, ctx->noAllocScratch);
    uint32_t dropped = 0;
    if (!check(hipGetLastError(), "allocation-free ray query launch") ||
        !check(hipStreamSynchronize(s), "allocation-free ray query") ||
        !check(hipMemcpy(&dropped, ctx->noAllocScratch, sizeof(dropped), hipMemcpyDeviceToHost), "read counter"))
        return -4;
    if (dropped) {
        setError("traversal stack overflow");
        return -5;
    }
    return 0;
}

extern "C" int mrt_trace_rays(mrt_context* ctx, const float* rays, size_t count, uint32_t* outMesh, uint32_t* outPrim,
                              float* outT, float* outBary) {
    if (!ctx || !rays || !outMesh || !outPrim) return -1;
    std::vector<float4> res;
    int rc = runQuery(ctx, rays, count, 0u, res);
    if (rc) return rc;
    std::vector<uint32_t> triMesh(ctx->triMesh.count), triBase(ctx->meshTriBase.count);
    if (!triMesh.empty()) hipMemcpy(triMesh.data(), ctx->triMesh.ptr, triMesh.size() * 4, hipMemcpyDeviceToHost);
    if (!triBase.empty()) hipMemcpy(triBase.data(), ctx->meshTriBase.ptr, triBase.size() * 4, hipMemcpyDeviceToHost);
    for (size_t i = 0; i < count; i++) {
        uint32_t tri = hostBits(res[i].y);
        if (tri == kNoHit || tri >= triMesh.size()) {
            outMesh[i] = VKRT_INVALID_INDEX;
            outPrim[i] = VKRT_INVALID_INDEX;
        } else {
            outMesh[i] = triMesh[tri];
            outPrim[i] = tri - triBase[triMesh[tri]];
        }
        if (outT) outT[i] = tri == kNoHit ? 0.0f : res[i].x;
        if (outBary) {
            outBary[2 * i] = tri == kNoHit ? 0.0f : res[i].z;
            outBary[2 * i + 1] = tri == kNoHit ? 0.0f : res[i].w;
        }
    }
    return 0;
}

extern "C" int mrt_trace_shadow_rays(mrt_context* ctx, const float* rays, size_t count, uint32_t* outVisibility) {
    if (!ctx || !rays || !outVisibility) return -1;
    std::vector<float4> res;
    int rc = runQuery(ctx, rays, count, 1u, res);
    if (rc) return rc;
    for (size_t i = 0; i < count; i++) outVisibility[i] = hostBits(res[i].x);
    return 0;
}

extern "C" int mrt_primary_hits(mrt_context* ctx, const SceneData* scene, const mrt_render_params* prm,
                                uint32_t sampleIndex, uint32_t* outMesh, uint32_t* outPrim) {
    if (!ctx || !scene || !prm || !outMesh || !outPrim) return -1;
    if (!check(hipSetDevice(ctx->device), "hipSetDevice")) return -2;
    if (!ensureTraversalScratch(ctx)) return -3;
    size_t n = (size_t)scene->viewportRect[2] * scene->viewportRect[3];
    if (n == 0) return 0;
    uint32_t *dMesh = nullptr, *dPrim = nullptr;
    if (!check(hipMalloc(&dMesh, n * 4), "hipMalloc") || !check(hipMalloc(&dPrim, n * 4), "hipMalloc")) {
        if (dMesh) hipFree(dMesh);
        return -3;
    }
    hipLaunchKernelGGL(kPrimary, dim3(ctx->traceGrid), dim3(kTraceBlock), 0, ctx->stream, ctx->view(), *scene,
                       prm->rngPolicy, prm->rngSeed, sampleIndex, dMesh, dPrim, ctx->traversalStack.ptr);
    int rc = 0;
    if (!check(hipStreamSynchronize(ctx->stream), "primary hits") ||
        !check(hipMemcpy(outMesh, dMesh, n * 4, hipMemcpyDeviceToHost), "read") ||
        !check(hipMemcpy(outPrim, dPrim, n * 4, hipMemcpyDeviceToHost), "read"))
        rc = -4;
    hipFree(dMesh);
    hipFree(dPrim);
    return rc;
}
