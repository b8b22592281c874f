// Layered principled BSDF on the GPU (src/shaders/bsdf/): microfiber sheen
// over clearcoat over a metal / dielectric / diffuse+subsurface base,
// one-sample lobe selection, RGB, single-wavelength and four-wavelength (hero)
// evaluation.
#pragma once

namespace mrt {

// Microfiber sheen LTC fit (Zeltner, Burley, Chiang 2022, as tabulated by
// Blender Cycles): [layer][roughness row][cos-theta column], 32 x 32 per
// layer; layers are the LTC transform a, transform b and directional albedo
// (src/shaders/bsdf/data/sheen_ltc.slang). 12 KB, read through the caches.
constexpr uint32_t kSheenLtcSize = 32u;
static __constant__ float kSheenLtc[3u * kSheenLtcSize * kSheenLtcSize] = {
#include "../sheen_ltc_table.inc"
};

// src/shaders/bsdf/base/material.slang:4-48 (fields the lobes read)
struct BMat {
    F3 baseColor;
    float roughness;
    F3 eta;
    float metallic;
    F3 k;
    float anisotropic, specular, specularTint, abbeNumber;
    float clearcoat, clearcoatGloss, ior, diffuseRoughness, transmission, subsurface;
    float absorptionCoefficient;
    F3 attenuationColor;
    F3 emission;  // emissionColor * emissionLuminance after the texture stage
    F3 sheenColor;  // saturate(tint) * saturate(weight), src/shaders/bsdf/base/material.slang:63-73
    float sheenRoughness;
};

// Material fetch plus applySurfaceTextures with no texture bound
// (src/shaders/material/textures.slang:123-141): base colour is modulated by
// the interpolated vertex colour; emission is renormalised by its maximum.
// The texture factors default to an empty slot's fallback sample (1).
DEV BMat loadMaterial(const Material* mats, uint32_t idx, F3 vertexColor, F3 texBase = F3{1.0f, 1.0f, 1.0f},
                      float texRough = 1.0f, float texMetal = 1.0f, F3 texEmission = F3{1.0f, 1.0f, 1.0f}) {
    const Material& m = mats[idx];
    BMat b;
    b.baseColor = F3{m.baseColor[0] * (texBase.x * vertexColor.x), m.baseColor[1] * (texBase.y * vertexColor.y),
                     m.baseColor[2] * (texBase.z * vertexColor.z)};
    b.roughness = sat(m.roughness * texRough);
    b.eta = F3{m.eta[0], m.eta[1], m.eta[2]};
    b.metallic = sat(m.metallic * texMetal);
    b.k = F3{m.k[0], m.k[1], m.k[2]};
    b.anisotropic = m.anisotropic;
    b.specular = m.specular;
    b.specularTint = m.specularTint;
    b.abbeNumber = m.abbeNumber;
    b.clearcoat = m.clearcoat;
    b.clearcoatGloss = m.clearcoatGloss;
    b.ior = m.ior;
    b.diffuseRoughness = m.diffuseRoughness;
    b.transmission = m.transmission;
    b.subsurface = m.subsurface;
    b.absorptionCoefficient = m.absorptionCoefficient;
    b.attenuationColor = F3{m.attenuationColor[0], m.attenuationColor[1], m.attenuationColor[2]};
    b.sheenColor = sat(F3{m.sheenTintWeight[0], m.sheenTintWeight[1], m.sheenTintWeight[2]}) * sat(m.sheenTintWeight[3]);
    b.sheenRoughness = m.sheenRoughness;
    F3 em = F3{m.emissionColor[0], m.emissionColor[1], m.emissionColor[2]} * m.emissionLuminance;
    em = em * texEmission;
    float emax = maxc(em);
    if (emax > 0.0f) {
        F3 col = em / emax;
        b.emission = col * emax;
    } else {
        b.emission = f3(0.0f);
    }
    return b;
}

// applySurfaceTextures (textures.slang:123-141) for a reconstructed hit: the
// base-colour, metallic-roughness (g, b) and emissive samples scale the
// material; untextured materials take the plain path.
DEV BMat loadSurfaceMaterial(const SceneView& s, const Surface& sf) {
    if (!sf.textured) return loadMaterial(s.materials, sf.materialIndex, sf.vertexColor);
    const Material& m = s.materials[sf.materialIndex];
    const F4 one = f4(1.0f);
    const F4 bc = sampleMaterialSlot(s, m, VKRT_MATERIAL_TEXTURE_SLOT_BASE_COLOR, sf.tc, one);
    const F4 mr = sampleMaterialSlot(s, m, VKRT_MATERIAL_TEXTURE_SLOT_METALLIC_ROUGHNESS, sf.tc, one);
    const F4 em = sampleMaterialSlot(s, m, VKRT_MATERIAL_TEXTURE_SLOT_EMISSIVE, sf.tc, one);
    return loadMaterial(s.materials, sf.materialIndex, sf.vertexColor, F3{bc.x, bc.y, bc.z}, mr.y, mr.z,
                        F3{em.x, em.y, em.z});
}

DEV F3 diffuseColor(const BMat& m) { return m.baseColor * (1.0f - m.metallic); }
DEV F3 tintColor(F3 c) {
    float l = lum(c);
    return l <= 0.0f ? f3(1.0f) : c / l;
}
DEV F3 transmissionColor(const BMat& m) {
    return (m.transmission > 0.0f && m.metallic <= 0.0f) ? f3(1.0f) : m.baseColor;
}
DEV float pow5(float v) {
    float v2 = v * v;
    return v2 * v2 * v;
}
DEV float schlickW(float c) { return pow5(1.0f - sat(c)); }
DEV float sin2T(F3 w) { return sat(1.0f - w.z * w.z); }
DEV float cosHemiPdf(F3 w) { return w.z > 0.0f ? w.z * kInvPi : 0.0f; }

// src/shaders/bsdf/base/fresnel.slang:4-68
DEV float dielectricF0(float eta) {
    float n = eta - 1.0f, d = eta + 1.0f;
    float f = n / fmaxf(d, 1e-6f);
    return f * f;
}
DEV bool hasConductor(const BMat& m) { return m.k.x > 0.0f || m.k.y > 0.0f || m.k.z > 0.0f; }
DEV F3 specularF0(const BMat& m) {
    float d = dielectricF0(m.ior);
    float ss = m.specular / 0.5f;
    F3 tint = lerp(f3(1.0f), tintColor(m.baseColor), sat(m.specularTint));
    return sat(tint * (d * ss));
}
DEV F3 schlick(float c, F3 f0) {
    float w = schlickW(c);
    return f0 + (f3(1.0f) - f0) * w;
}
DEV float fresnelDielectric(float cosTheta, float eta) {
    float ci = clampf(cosTheta, -1.0f, 1.0f);
    if (ci < 0.0f) {
        eta = 1.0f / fmaxf(eta, 1e-6f);
        ci = -ci;
    }
    float s2i = fmaxf(1.0f - ci * ci, 0.0f);
    float s2t = s2i / fmaxf(eta * eta, 1e-6f);
    if (s2t >= 1.0f) return 1.0f;
    float ct = sqrtf(fmaxf(1.0f - s2t, 0.0f));
    float rpa = (eta * ci - ct) / fmaxf(eta * ci + ct, 1e-6f);
    float rpe = (ci - eta * ct) / fmaxf(ci + eta * ct, 1e-6f);
    return 0.5f * (rpa * rpa + rpe * rpe);
}
DEV float conductor1(float ci, float eta, float k) {
    float c2 = ci * ci;
    float s2 = fmaxf(1.0f - c2, 0.0f);
    float eta2 = eta * eta, k2 = k * k;
    float t0 = eta2 - k2 - s2;
    float a2b2 = sqrtf(fmaxf(t0 * t0 + 4.0f * eta2 * k2, 0.0f));
    float t1 = a2b2 + c2;
    float a = sqrtf(fmaxf(0.5f * (a2b2 + t0), 0.0f));
    float t2 = 2.0f * ci * a;
    float rs = (t1 - t2) / fmaxf(t1 + t2, 1e-6f);
    float t3 = c2 * a2b2 + s2 * s2;
    float t4 = t2 * s2;
    float rp = rs * ((t3 - t4) / fmaxf(t3 + t4, 1e-6f));
    return 0.5f * (rp + rs);
}
DEV F3 metalFresnel(const BMat& m, float c) {
    if (hasConductor(m)) {
        float ci = sat(c);
        F3 e = maxs(m.eta, 1e-3f), kk = maxs(m.k, 0.0f);
        return F3{conductor1(ci, e.x, kk.x), conductor1(ci, e.y, kk.y), conductor1(ci, e.z, kk.z)};
    }
    return schlick(c, m.baseColor);
}

// src/shaders/bsdf/base/interface.slang, medium.slang:35-61
DEV float interfaceIor(const BMat& m) { return fmaxf(m.ior, 1.0f); }
DEV float etaFor(const BMat& m, uint32_t ff, float l, uint32_t spectral) {
    float ior = interfaceIor(m);
    if (spectral) ior = dispersiveIor(ior, m.abbeNumber, l);
    return ff ? ior : 1.0f / ior;
}
// Whether eta depends on the wavelength: dispersiveIor returns max(ior, 1)
// for every wavelength when this is false, so the hero lanes of such a
// material share one eta and everything computed from it.
DEV bool dispersive(const BMat& m) {
    const float ior = interfaceIor(m);
    return !(ior <= 1.0f + 1e-4f || m.abbeNumber <= 0.0f);
}
DEV float refractionEtaFor(const BMat& m, uint32_t ff, float l, uint32_t spectral) {
    float ior = interfaceIor(m);
    if (spectral) ior = dispersiveIor(ior, m.abbeNumber, l);
    return ff ? 1.0f / ior : ior;
}
DEV bool hasAbsorption(const BMat& m) {
    return m.absorptionCoefficient > 0.0f &&
           (m.attenuationColor.x < 0.9999f || m.attenuationColor.y < 0.9999f || m.attenuationColor.z < 0.9999f);
}
DEV bool isRefractive(const BMat& m) { return m.transmission > 0.0f && interfaceIor(m) > 1.0f + 1e-4f; }
DEV bool dielectricIdentity(const BMat& m) { return interfaceIor(m) <= 1.0f + 1e-4f; }

// ----------------------------------------------------------------- GGX --
// src/shaders/bsdf/lobes/ggx.slang:16-227
struct GGX {
    float ax, ay, vndfK, projLen;
};
DEV GGX makeGGX(const BMat& m, F3 wo) {
    GGX g;
    float rough = fmaxf(m.roughness, 1e-3f);
    float aspect = sqrtf(fmaxf(1.0f - 0.9f * m.anisotropic, 1e-3f));
    float a = rough * rough;
    g.ax = fmaxf(a / aspect, 1e-3f);
    g.ay = fmaxf(a * aspect, 1e-3f);
    float am = sat(fminf(g.ax, g.ay));
    float s = 1.0f + sqrtf(wo.x * wo.x + wo.y * wo.y);
    float a2 = am * am, s2 = s * s;
    g.vndfK = (1.0f - a2) * s2 / fmaxf(s2 + a2 * wo.z * wo.z, 1e-6f);
    float sx = g.ax * wo.x, sy = g.ay * wo.y;
    g.projLen = sqrtf(sx * sx + sy * sy + wo.z * wo.z);
    return g;
}
DEV float ggxLambda(F3 w, const GGX& g) {
    float z2 = fmaxf(w.z * w.z, 1e-6f);
    float slope2 = (g.ax * g.ax * w.x * w.x + g.ay * g.ay * w.y * w.y) / z2;
    return 0.5f * (sqrtf(1.0f + slope2) - 1.0f);
}
DEV float ggxD(F3 m, const GGX& g) {
    if (m.z <= 0.0f) return 0.0f;
    float c = (m.x * m.x) / (g.ax * g.ax) + (m.y * m.y) / (g.ay * g.ay) + m.z * m.z;
    return 1.0f / (kPi * g.ax * g.ay * c * c);
}
DEV float ggxG(F3 wo, F3 wi, const GGX& g) { return 1.0f / (1.0f + ggxLambda(wo, g) + ggxLambda(wi, g)); }
DEV float ggxG1(F3 w, const GGX& g) { return 1.0f / (1.0f + ggxLambda(w, g)); }
DEV bool ggxHalf(F3 wo, F3 wi, F3& m, float& wom) {
    F3 h = wo + wi;
    float h2 = dot(h, h);
    if (h2 <= 1e-6f) {
        m = f3(0.0f);
        wom = 0.0f;
        return false;
    }
    m = h * rsq(h2);
    wom = dot(wo, m);
    return wom > 0.0f;
}
DEV float ggxVisiblePdf(F3 wo, F3 m, const GGX& g) {
    float wm = dot(wo, m);
    if (wo.z <= 0.0f || m.z <= 0.0f || wm <= 0.0f) return 0.0f;
    return ggxD(m, g) * ggxG1(wo, g) * wm / fmaxf(wo.z, 1e-6f);
}
DEV float ggxReflPdf(F3 wo, F3 m, const GGX& g) { return ggxVisiblePdf(wo, m, g) / fmaxf(4.0f * fabsf(dot(wo, m)), 1e-6f); }
DEV bool sampleVNDF(F3 wo, const GGX& g, float u0, float u1, F3& m) {
    if (g.projLen <= 1e-6f) return false;
    F3 ws = F3{wo.x * g.ax, wo.y * g.ay, wo.z} / g.projLen;
    float phi = 2.0f * kPi * u0;
    float b = g.vndfK * ws.z;
    float z = (1.0f - u1) * (1.0f + b) - b;
    float st = sqrtf(sat(1.0f - z * z));
    F3 wis = F3{st * cosf(phi), st * sinf(phi), z};
    F3 ms = ws + wis;
    if (dot(ms, ms) <= 1e-6f) return false;
    F3 um = F3{ms.x * g.ax, ms.y * g.ay, ms.z};
    float um2 = dot(um, um);
    if (um2 <= 1e-6f) return false;
    m = um * rsq(um2);
    return m.z > 0.0f;
}

struct Eval {
    F3 value;
    float pdf;
};

// ------------------------------------------------------ diffuse family --
// src/shaders/bsdf/lobes/diffuse.slang:13-53, subsurface.slang:11-31
DEV float cosPhiDiff(F3 wi, F3 wo) {
    float a = sin2T(wi), b = sin2T(wo);
    if (a <= 0.0f || b <= 0.0f) return 0.0f;
    float inv = rsq(a * b);
    return clampf((wi.x * wo.x + wi.y * wo.y) * inv, -1.0f, 1.0f);
}
DEV Eval evalDiffuse(F3 col, const BMat& m, F3 wo, F3 wi) {
    Eval e{f3(0.0f), 0.0f};
    if (m.diffuseRoughness <= 0.0f) {
        e.value = wi.z > 0.0f ? col * kInvPi : f3(0.0f);
        e.pdf = cosHemiPdf(wi);
        return e;
    }
    if (wo.z <= 0.0f || wi.z <= 0.0f) return e;
    float sigma = sat(m.diffuseRoughness) * (0.5f * kPi);
    float s2 = sigma * sigma;
    float A = 1.0f - s2 / (2.0f * (s2 + 0.33f));
    float B = 0.45f * s2 / (s2 + 0.09f);
    float sti = sqrtf(sin2T(wi)), sto = sqrtf(sin2T(wo));
    float mc = fmaxf(0.0f, cosPhiDiff(wi, wo));
    float sa, tb;
    if (fabsf(wi.z) > fabsf(wo.z)) {
        sa = sto;
        tb = sti / fmaxf(fabsf(wi.z), 1e-6f);
    } else {
        sa = sti;
        tb = sto / fmaxf(fabsf(wo.z), 1e-6f);
    }
    e.value = col * (kInvPi * (A + B * mc * sa * tb));
    e.pdf = cosHemiPdf(wi);
    return e;
}
DEV Eval evalSubsurface(F3 col, float rough, F3 wo, F3 wi) {
    Eval e{f3(0.0f), 0.0f};
    if (wo.z <= 0.0f || wi.z <= 0.0f) return e;
    F3 h = safeNormalize(wo + wi);
    float wih = sat(dot(wi, h));
    float f90 = wih * wih * rough;
    float fi = lerpf(1.0f, f90, schlickW(wi.z));
    float fo = lerpf(1.0f, f90, schlickW(wo.z));
    float f = fi * fo;
    float ss = 1.25f * (f * (1.0f / fmaxf(wi.z + wo.z, 1e-6f) - 0.5f) + 0.5f);
    e.value = col * (kInvPi * ss);
    e.pdf = cosHemiPdf(wi);
    return e;
}

// ------------------------------------------------------------ clearcoat --
// src/shaders/bsdf/lobes/clearcoat.slang:14-97
DEV float ccAlpha(const BMat& m) { return fmaxf(lerpf(0.1f, 0.001f, sat(m.clearcoatGloss)), 1e-3f); }
DEV float ccG1(float c, float a) {
    float a2 = a * a;
    float c2 = fmaxf(c * c, 1e-6f);
    float t2 = fmaxf(1.0f - c2, 0.0f) / c2;
    return 2.0f / (1.0f + sqrtf(1.0f + a2 * t2));
}
DEV float ccD(float cm, float a) {
    float a2 = a * a;
    if (a2 >= 1.0f - 1e-6f) return kInvPi;
    float den = kPi * logf(a2) * (1.0f + (a2 - 1.0f) * cm * cm);
    return (a2 - 1.0f) / fminf(den, -1e-6f);
}
DEV Eval evalClearcoat(float w, F3 wo, F3 wi, float alpha) {
    Eval e{f3(0.0f), 0.0f};
    if (w <= 0.0f || wo.z <= 0.0f || wi.z <= 0.0f) return e;
    F3 h = safeNormalize(wo + wi);
    float woh = sat(dot(wo, h));
    if (woh <= 0.0f) return e;
    float D = ccD(sat(h.z), alpha);
    float F = lerpf(0.04f, 1.0f, schlickW(woh));
    float G = ccG1(wo.z, 0.25f) * ccG1(wi.z, 0.25f);
    float v = 0.25f * w * D * F * G / fmaxf(4.0f * wo.z * wi.z, 1e-6f);
    e.value = f3(v);
    e.pdf = D * sat(h.z) / fmaxf(4.0f * woh, 1e-6f);
    return e;
}

// ----------------------------------------------------------- dielectric --
// src/shaders/bsdf/lobes/dielectric/{interface,eval_rgb,eval_spectral,sample}.slang
DEV float transmissionProb(const BMat& m, float F) {
    float tw = m.transmission * (1.0f - F);
    float tot = F + tw;
    return tot <= 0.0f ? 0.0f : tw / tot;
}
DEV bool transHalf(F3 wo, F3 wi, float etap, F3& wm, float& woD, float& wiD, float& den) {
    wm = f3(0.0f);
    woD = wiD = den = 0.0f;
    if (wo.z <= 0.0f || wi.z >= 0.0f) return false;
    F3 h = wo + wi * etap;
    if (dot(h, h) <= 1e-6f) return false;
    wm = safeNormalize(h);
    if (wm.z < 0.0f) wm = -wm;
    woD = dot(wo, wm);
    wiD = dot(wi, wm);
    den = wiD + woD / etap;
    return woD > 0.0f && wiD < 0.0f && fabsf(den) > 1e-6f;
}
// One transmission lane: value (given the lane's transmission colour) and pdf.
// The transmission lobe of one wavelength lane in two parts, so that lanes
// with the same eta share the first: its value is coatAtt * tc * x, or
// coatAtt * tc / x for an index-matched interface, or 0 (kind 0, 1, 2).
struct TransPart {
    float x, pdf;
    uint32_t kind;
};
DEV TransPart transPart(const BMat& m, F3 wo, F3 wi, uint32_t ff, const GGX& g, float l, uint32_t spectral) {
    TransPart t{0.0f, 0.0f, 0u};
    if (m.transmission <= 0.0f || wo.z <= 0.0f || wi.z >= 0.0f) return t;
    if (dielectricIdentity(m)) {
        if (dot(wi, -wo) < 0.9999f) return t;
        t.pdf = 1.0f;
        t.x = fmaxf(fabsf(wi.z), 1e-6f);
        t.kind = 2u;
        return t;
    }
    float etap = etaFor(m, ff, l, spectral);
    F3 wm;
    float woD, wiD, den;
    if (!transHalf(wo, wi, etap, wm, woD, wiD, den)) return t;
    float F = fresnelDielectric(woD, etap);
    float tp = transmissionProb(m, F);
    if (tp <= 0.0f) return t;
    float D = ggxD(wm, g), G = ggxG(wo, wi, g);
    float den2 = den * den;
    float ts = 1.0f / fmaxf(etap * etap, 1e-6f);
    float tt = fabsf((wiD * woD) / fmaxf(fabsf(wi.z) * wo.z * den2, 1e-6f));
    float dwm = fabsf(wiD) / fmaxf(den2, 1e-6f);
    t.pdf = tp * ggxVisiblePdf(wo, wm, g) * dwm;
    t.x = (1.0f - F) * D * G * tt * ts;
    t.kind = 1u;
    return t;
}
DEV float transValue(const TransPart& t, float coatAtt, float tc) {
    return t.kind == 1u ? coatAtt * tc * t.x : (t.kind == 2u ? coatAtt * tc / t.x : 0.0f);
}
DEV float evalTransLane(const BMat& m, F3 wo, F3 wi, uint32_t ff, const GGX& g, float coatAtt, float tc, float l,
                        uint32_t spectral, float& pdf) {
    const TransPart t = transPart(m, wo, wi, ff, g, l, spectral);
    pdf = t.pdf;
    return transValue(t, coatAtt, tc);
}

// --------------------------------------------------------- principled --
struct Weights {
    float sheen, coat, metal, dielectric, diffuse, subsurface;
};

// ---------------------------------------------------------------- sheen --
// src/shaders/bsdf/lobes/sheen.slang:22-113. The lobe is a linearly
// transformed cosine in a frame whose x axis is wo projected onto the surface.
DEV float sheenLtc(float cosTheta, float roughness, uint32_t layer) {
    const float last = (float)(kSheenLtcSize - 1u);
    const float ux = sat(cosTheta) * last, uy = sat(roughness) * last;
    const uint32_t x0 = min((uint32_t)ux, kSheenLtcSize - 1u), y0 = min((uint32_t)uy, kSheenLtcSize - 1u);
    const uint32_t x1 = min(x0 + 1u, kSheenLtcSize - 1u), y1 = min(y0 + 1u, kSheenLtcSize - 1u);
    const float tx = ux - floorf(ux), ty = uy - floorf(uy);
    const float* t = kSheenLtc + layer * kSheenLtcSize * kSheenLtcSize;
    const float e00 = t[y0 * kSheenLtcSize + x0], e10 = t[y0 * kSheenLtcSize + x1];
    const float e01 = t[y1 * kSheenLtcSize + x0], e11 = t[y1 * kSheenLtcSize + x1];
    return lerpf(lerpf(e00, e10, tx), lerpf(e01, e11, tx), ty);
}

struct SheenLtc {
    Basis frame;
    float a, b, albedo;
};

DEV SheenLtc makeSheenLtc(F3 wo, float sheenRoughness) {
    SheenLtc p;
    p.frame = makeBasis(F3{0.0f, 0.0f, 1.0f}, F4{wo.x, wo.y, wo.z, 1.0f});
    p.a = p.b = p.albedo = 0.0f;
    if (wo.z <= 0.0f) return p;
    const float r = clampf(sheenRoughness, 1e-3f, 1.0f);
    p.a = sheenLtc(wo.z, r, 0u);
    p.b = sheenLtc(wo.z, r, 1u);
    p.albedo = sheenLtc(wo.z, r, 2u);
    return p;
}

// LTC density of a direction given in the sheen frame (also the sampling pdf)
DEV float sheenDensity(F3 lw, const SheenLtc& p) {
    const float z = fmaxf(lw.z, 0.0f);
    if (z <= 0.0f || fabsf(p.a) < 1e-5f || p.albedo < 1e-5f) return 0.0f;
    const float axbz = p.a * lw.x + p.b * lw.z, ay = p.a * lw.y;
    const float len2 = axbz * axbz + ay * ay + lw.z * lw.z;
    if (len2 <= 1e-6f) return 0.0f;
    const float scale = p.a / len2;
    return kInvPi * z * scale * scale;
}

DEV float sheenAlbedo(float cosTheta, float sheenRoughness) {
    return sheenLtc(cosTheta, clampf(sheenRoughness, 1e-3f, 1.0f), 2u);
}
DEV float sheenLayerAtt(float weight, float cosTheta, float sheenRoughness) {
    return sat(1.0f - weight * sheenAlbedo(cosTheta, sheenRoughness));
}
// src/shaders/bsdf/principled/attenuation.slang:25-28
DEV float sheenAtt(F3 sheenColor, F3 w, float sheenRoughness) {
    return sheenLayerAtt(maxc(sheenColor), w.z, sheenRoughness);
}

// evalSheen with a unit colour: value = albedo * density, pdf = density
DEV Eval evalSheenLobe(float sheenRoughness, F3 wo, F3 wi) {
    Eval e{f3(0.0f), 0.0f};
    if (wo.z <= 0.0f || wi.z <= 0.0f) return e;
    const SheenLtc p = makeSheenLtc(wo, sheenRoughness);
    const float d = sheenDensity(toLocal(wi, p.frame), p);
    e.value = f3(p.albedo * d);
    e.pdf = d;
    return e;
}

struct BState {
    BMat m;
    F3 wo;
    GGX g;
    float ccAlpha;
    uint32_t ff;
    float l;
    uint32_t spectral;
    Weights w;
};

// src/shaders/bsdf/principled/attenuation.slang:4-66
DEV float coatAtt(const BMat& m, F3 w) {
    if (m.clearcoat <= 0.0f) return 1.0f;
    if (fabsf(w.z) <= 0.0f) return 0.0f;
    float f = 0.25f * m.clearcoat * lerpf(0.04f, 1.0f, schlickW(fabsf(fabsf(w.z))));
    return 1.0f - sat(f);
}
DEV F3 dielectricAtt(const BMat& m, F3 w) {
    if (fabsf(w.z) <= 0.0f) return sat(f3(1.0f) - f3(1.0f));
    float cr = sat(m.roughness), cc = sat(fabsf(w.z));
    float rx = cr * -1.0f + 1.0f, ry = cr * -0.0275f + 0.0425f, rz = cr * -0.572f + 1.04f, rw = cr * 0.022f + -0.04f;
    float a004 = fminf(rx * rx, exp2f(-9.28f * cc)) * rx + ry;
    float abx = sat(-1.04f * a004 + rz), aby = sat(1.04f * a004 + rw);
    F3 albedo = sat(specularF0(m) * abx + f3(aby));
    return sat(f3(1.0f) - albedo);
}

// src/shaders/bsdf/principled/weights.slang:4-62. SH = false compiles the
// sheen layer out (weight 0, attenuation 1) for scenes without sheen, which
// keeps the shading kernel inside its register budget.
template <bool SH = true>
DEV Weights makeWeights(const BMat& m, F3 wo, uint32_t ff) {
    Weights w;
    F3 dc = diffuseColor(m);
    float nonMetal = 1.0f - m.metallic;
    float baseW = fmaxf(lum(m.baseColor), 1e-3f);
    float diffW = fmaxf(lum(dc), 1e-3f);
    float baseScatter = nonMetal * (1.0f - m.transmission);
    float transW = m.transmission * fmaxf(lum(transmissionColor(m)), 1e-3f);
    float dielW = nonMetal * fmaxf(sat(lum(specularF0(m))) + transW, 0.0f);
    float total;
    w.sheen = SH ? lum(m.sheenColor) : 0.0f;
    if (m.transmission > 0.0f && ff == 0u) {
        w.sheen = 0.0f;
        w.coat = 0.0f;
        w.metal = m.metallic * baseW;
        w.dielectric = dielW;
        w.diffuse = 0.0f;
        w.subsurface = 0.0f;
    } else {
        const float vs = SH ? sheenAtt(m.sheenColor, wo, m.sheenRoughness) : 1.0f;
        float vc = coatAtt(m, wo);
        float da = lum(dielectricAtt(m, wo));
        w.coat = vs * m.clearcoat * 0.25f;
        w.metal = vs * vc * m.metallic * baseW;
        w.dielectric = vs * vc * dielW;
        w.diffuse = vs * vc * da * baseScatter * (1.0f - m.subsurface) * diffW;
        w.subsurface = vs * vc * da * baseScatter * m.subsurface * diffW;
    }
    total = w.sheen + w.coat + w.metal + w.dielectric + w.diffuse + w.subsurface;
    if (total <= 0.0f) {
        w = Weights{0.0f, 0.0f, 0.0f, 0.0f, 1.0f, 0.0f};
        return w;
    }
    w.sheen /= total;
    w.coat /= total;
    w.metal /= total;
    w.dielectric /= total;
    w.diffuse /= total;
    w.subsurface /= total;
    return w;
}

template <bool SH = true>
DEV BState makeState(const BMat& m, F3 wo, uint32_t ff, float l, uint32_t spectral) {
    BState s;
    s.m = m;
    s.wo = wo;
    s.ff = ff;
    s.l = l;
    s.spectral = spectral;
    s.g = makeGGX(m, wo);
    s.ccAlpha = ccAlpha(m);
    s.w = makeWeights<SH>(m, wo, ff);
    return s;
}
DEV bool interior(const BState& s) { return s.m.transmission > 0.0f && s.ff == 0u; }

// src/shaders/bsdf/principled/eval_rgb.slang:4-140 (RGB and single wavelength)
template <bool SH = true>
DEV Eval evalScalar(const Spec& sp, const BState& s, F3 wi) {
    Eval e{f3(0.0f), 0.0f};
    const BMat& m = s.m;
    if (wi.z > 0.0f) {
        bool inner = interior(s);
        F3 coatV = f3(0.0f), metalV = f3(0.0f), dielV = f3(0.0f), subV = f3(0.0f), sheenV = f3(0.0f);
        F3 sheenColor = f3(0.0f);
        if constexpr (SH) {
            if (s.w.sheen > 0.0f || !inner) {
                sheenColor = m.sheenColor;
                if (s.spectral) sheenColor = f3(spec1(sp, sat(sheenColor), s.l));
            }
            if (s.w.sheen > 0.0f) {
                Eval sh = evalSheenLobe(m.sheenRoughness, s.wo, wi);
                sheenV = sheenColor * sh.value.x;
                e.pdf += s.w.sheen * sh.pdf;
            }
        }
        if (s.w.coat > 0.0f) {
            Eval c = evalClearcoat(m.clearcoat, s.wo, wi, s.ccAlpha);
            coatV = c.value;
            e.pdf += s.w.coat * c.pdf;
        }
        if (s.w.metal > 0.0f && s.wo.z > 0.0f) {
            F3 h;
            float wom;
            if (ggxHalf(s.wo, wi, h, wom)) {
                float D = ggxD(h, s.g), G = ggxG(s.wo, wi, s.g);
                F3 F = metalFresnel(m, wom);
                if (s.spectral) F = f3(spec1(sp, sat(F), s.l));
                metalV = F * (D * G / fmaxf(4.0f * s.wo.z * wi.z, 1e-6f));
                e.pdf += s.w.metal * ggxReflPdf(s.wo, h, s.g);
            }
        }
        if (s.w.dielectric > 0.0f && !dielectricIdentity(m) && s.wo.z > 0.0f) {
            F3 h;
            float wom;
            if (ggxHalf(s.wo, wi, h, wom)) {
                float F = fresnelDielectric(wom, etaFor(m, s.ff, s.l, s.spectral));
                float rp = 1.0f - transmissionProb(m, F);
                if (rp > 0.0f) {
                    float D = ggxD(h, s.g), G = ggxG(s.wo, wi, s.g);
                    F3 rc;
                    if (m.transmission > 0.0f && m.metallic <= 0.0f) {
                        rc = f3(F);
                    } else {
                        rc = schlick(wom, specularF0(m));
                        if (s.spectral) rc = f3(spec1(sp, sat(rc), s.l));
                    }
                    dielV = rc * (D * G / fmaxf(4.0f * s.wo.z * wi.z, 1e-6f));
                    e.pdf += s.w.dielectric * (rp * ggxReflPdf(s.wo, h, s.g));
                }
            }
        }
        if (s.w.diffuse > 0.0f || s.w.subsurface > 0.0f) {
            F3 dc = diffuseColor(m);
            if (s.spectral) dc = f3(spec1(sp, sat(dc), s.l));
            if (s.w.diffuse > 0.0f) {
                Eval d = evalDiffuse(dc * ((1.0f - m.transmission) * (1.0f - m.subsurface)), m, s.wo, wi);
                subV = subV + d.value;
                e.pdf += s.w.diffuse * d.pdf;
            }
            if (s.w.subsurface > 0.0f) {
                Eval d = evalSubsurface(dc * (1.0f - m.transmission) * m.subsurface, m.diffuseRoughness, s.wo, wi);
                subV = subV + d.value;
                e.pdf += s.w.subsurface * d.pdf;
            }
        }
        float nonMetal = 1.0f - m.metallic;
        if (inner) {
            e.value = metalV * m.metallic + dielV * nonMetal;
            return e;
        }
        float vs = SH ? sheenAtt(sheenColor, s.wo, m.sheenRoughness) : 1.0f;
        float vc = coatAtt(m, s.wo);
        F3 vd = dielectricAtt(m, s.wo);
        F3 base = coatV + (metalV * m.metallic + (dielV + vd * subV) * nonMetal) * (vc * coatAtt(m, wi));
        e.value = sheenV + base * vs;
        return e;
    }
    // src/shaders/bsdf/principled/attenuation.slang:57-66: sheen and coat of the
    // material's own (RGB) sheen colour over the transmitted lobe
    float ca = interior(s) ? 1.0f : (SH ? sheenAtt(m.sheenColor, s.wo, m.sheenRoughness) : 1.0f) * coatAtt(m, s.wo);
    F3 tc = transmissionColor(m) * ((1.0f - m.metallic) * m.transmission);
    if (s.spectral) tc = f3(spec1(sp, sat(tc), s.l));
    float pdf;
    float vx = evalTransLane(m, s.wo, wi, s.ff, s.g, ca, tc.x, s.l, s.spectral, pdf);
    if (pdf <= 0.0f && vx == 0.0f) {
        e.pdf = pdf * s.w.dielectric;
        return e;
    }
    float p2;
    float vy = tc.y == tc.x ? vx : evalTransLane(m, s.wo, wi, s.ff, s.g, ca, tc.y, s.l, s.spectral, p2);
    float vz = tc.z == tc.x ? vx : evalTransLane(m, s.wo, wi, s.ff, s.g, ca, tc.z, s.l, s.spectral, p2);
    e.value = F3{vx, vy, vz};
    e.pdf = pdf * s.w.dielectric;
    return e;
}

// src/shaders/bsdf/principled/eval_spectral.slang:4-115 (four hero lanes)
template <bool SH = true>
DEV F4 evalHero(const Spec& sp, const BState& s, F3 wi, F4 l, F4& tp) {
    const BMat& m = s.m;
    tp = f4(0.0f);
    if (wi.z > 0.0f) {
        bool inner = interior(s);
        float coatV = 0.0f;
        F4 metalV = f4(0.0f), dielV = f4(0.0f), subV = f4(0.0f), sheenV = f4(0.0f), sheenColor = f4(0.0f);
        if constexpr (SH) {
            if (s.w.sheen > 0.0f || !inner) sheenColor = spec4(sp, sat(m.sheenColor), l);
            if (s.w.sheen > 0.0f) {
                Eval sh = evalSheenLobe(m.sheenRoughness, s.wo, wi);
                sheenV = sheenColor * sh.value.x;
                tp = tp + f4(s.w.sheen * sh.pdf);
            }
        }
        if (s.w.coat > 0.0f) {
            Eval c = evalClearcoat(m.clearcoat, s.wo, wi, s.ccAlpha);
            coatV = c.value.x;
            tp = tp + f4(s.w.coat * c.pdf);
        }
        if (s.w.metal > 0.0f && s.wo.z > 0.0f) {
            F3 h;
            float wom;
            if (ggxHalf(s.wo, wi, h, wom)) {
                float D = ggxD(h, s.g), G = ggxG(s.wo, wi, s.g);
                float mp = ggxReflPdf(s.wo, h, s.g);
                metalV = spec4(sp, sat(metalFresnel(m, wom)), l) * (D * G / fmaxf(4.0f * s.wo.z * wi.z, 1e-6f));
                tp = tp + f4(s.w.metal * mp);
            }
        }
        if (s.w.dielectric > 0.0f) {
            F4 dp = f4(0.0f);
            if (!dielectricIdentity(m) && s.wo.z > 0.0f) {
                F3 h;
                float wom;
                if (ggxHalf(s.wo, wi, h, wom)) {
                    F4 F, rp;
                    if (dispersive(m)) {
                        F = F4{fresnelDielectric(wom, etaFor(m, s.ff, l.x, 1u)), fresnelDielectric(wom, etaFor(m, s.ff, l.y, 1u)),
                               fresnelDielectric(wom, etaFor(m, s.ff, l.z, 1u)), fresnelDielectric(wom, etaFor(m, s.ff, l.w, 1u))};
                        rp = F4{1.0f - transmissionProb(m, F.x), 1.0f - transmissionProb(m, F.y),
                                1.0f - transmissionProb(m, F.z), 1.0f - transmissionProb(m, F.w)};
                    } else {  // one eta for every lane
                        F = f4(fresnelDielectric(wom, etaFor(m, s.ff, l.x, 1u)));
                        rp = f4(1.0f - transmissionProb(m, F.x));
                    }
                    float D = ggxD(h, s.g), G = ggxG(s.wo, wi, s.g);
                    float rbp = ggxReflPdf(s.wo, h, s.g);
                    dp = rp * rbp;
                    if (anyPos(dp)) {
                        F4 rc;
                        if (m.transmission > 0.0f && m.metallic <= 0.0f)
                            rc = F;
                        else
                            rc = spec4(sp, sat(schlick(wom, specularF0(m))), l);
                        dielV = rc * (D * G / fmaxf(4.0f * s.wo.z * wi.z, 1e-6f));
                    } else {
                        dp = f4(0.0f);
                    }
                }
            }
            tp = tp + dp * s.w.dielectric;
        }
        if (s.w.diffuse > 0.0f || s.w.subsurface > 0.0f) {
            F4 dc = spec4(sp, sat(diffuseColor(m)), l);
            if (s.w.diffuse > 0.0f) {
                Eval d = evalDiffuse(f3((1.0f - m.transmission) * (1.0f - m.subsurface)), m, s.wo, wi);
                subV = subV + dc * d.value.x;
                tp = tp + f4(s.w.diffuse * d.pdf);
            }
            if (s.w.subsurface > 0.0f) {
                Eval d = evalSubsurface(f3((1.0f - m.transmission) * m.subsurface), m.diffuseRoughness, s.wo, wi);
                subV = subV + dc * d.value.x;
                tp = tp + f4(s.w.subsurface * d.pdf);
            }
        }
        float nonMetal = 1.0f - m.metallic;
        if (inner) return metalV * m.metallic + dielV * nonMetal;
        float vc = coatAtt(m, s.wo);
        F4 vs = f4(1.0f);
        if constexpr (SH) vs = sat(f4(1.0f) - sheenColor * sheenAlbedo(s.wo.z, m.sheenRoughness));
        F4 vd = spec4(sp, sat(dielectricAtt(m, s.wo)), l);
        float ra = vc * coatAtt(m, wi);
        F4 base = f4(coatV) + (metalV * m.metallic + (dielV + vd * subV) * nonMetal) * ra;
        return sheenV + vs * base;
    }
    float ca = interior(s) ? 1.0f : (SH ? sheenAtt(m.sheenColor, s.wo, m.sheenRoughness) : 1.0f) * coatAtt(m, s.wo);
    F4 tc = spec4(sp, sat(transmissionColor(m)), l) * ((1.0f - m.metallic) * m.transmission);
    F4 v;
    if (dispersive(m)) {
        v.x = evalTransLane(m, s.wo, wi, s.ff, s.g, ca, tc.x, l.x, 1u, tp.x);
        v.y = evalTransLane(m, s.wo, wi, s.ff, s.g, ca, tc.y, l.y, 1u, tp.y);
        v.z = evalTransLane(m, s.wo, wi, s.ff, s.g, ca, tc.z, l.z, 1u, tp.z);
        v.w = evalTransLane(m, s.wo, wi, s.ff, s.g, ca, tc.w, l.w, 1u, tp.w);
    } else {  // one eta for every lane: one evaluation of everything but the lane's colour
        const TransPart t = transPart(m, s.wo, wi, s.ff, s.g, l.x, 1u);
        v = F4{transValue(t, ca, tc.x), transValue(t, ca, tc.y), transValue(t, ca, tc.z), transValue(t, ca, tc.w)};
        tp = f4(t.pdf);
    }
    tp = tp * s.w.dielectric;
    return v;
}

// src/shaders/bsdf/principled/sample.slang:4-51 with the lobe samplers
template <bool SH = true>
DEV bool sampleLobe(const BState& s, Rng& rng, F3& wi, uint32_t& isT) {
    wi = f3(0.0f);
    isT = 0u;
    const BMat& m = s.m;
    float sel = rnd(rng);
    if (SH && sel < s.w.sheen) {
        // sheen.slang:99-113: a cosine-distributed disk sample through the LTC transform
        const SheenLtc p = makeSheenLtc(s.wo, m.sheenRoughness);
        if (fabsf(p.a) < 1e-5f || p.albedo < 1e-5f) return false;
        const float r = sqrtf(rnd(rng));
        const float phi = 2.0f * kPi * rnd(rng);
        const float dx = r * cosf(phi), dy = r * sinf(phi);
        const float dz = sqrtf(fmaxf(1.0f - (dx * dx + dy * dy), 0.0f));
        wi = toWorld(normalize(F3{dx - dz * p.b, dy, dz * p.a}), p.frame);
        return wi.z > 0.0f;
    }
    sel -= s.w.sheen;
    if (sel < s.w.coat) {
        float u1 = rnd(rng);
        float u2 = rnd(rng);
        float a2 = s.ccAlpha * s.ccAlpha;
        float cm;
        if (a2 >= 1.0f - 1e-6f) {
            cm = sqrtf(fmaxf(0.0f, 1.0f - u1));
        } else {
            float ex = powf(a2, 1.0f - u1);
            cm = sqrtf(sat((1.0f - ex) / fmaxf(1.0f - a2, 1e-6f)));
        }
        float sm = sqrtf(sat(1.0f - cm * cm));
        float phi = 2.0f * kPi * u2;
        F3 h = F3{sm * cosf(phi), sm * sinf(phi), cm};
        float woh = dot(s.wo, h);
        if (woh <= 0.0f) return false;
        wi = h * (2.0f * woh) - s.wo;
        return wi.z > 0.0f;
    }
    sel -= s.w.coat;
    if (sel < s.w.metal) {
        float u0 = rnd(rng);
        float u1 = rnd(rng);
        F3 h;
        if (!sampleVNDF(s.wo, s.g, u0, u1, h)) return false;
        float woh = dot(s.wo, h);
        if (woh <= 0.0f) return false;
        wi = h * (2.0f * woh) - s.wo;
        return wi.z > 0.0f;
    }
    sel -= s.w.metal;
    if (sel < s.w.dielectric) {
        if (s.wo.z <= 0.0f) return false;
        if (dielectricIdentity(m)) {
            wi = -s.wo;
            isT = m.transmission > 0.0f ? 1u : 0u;
            return m.transmission > 0.0f;
        }
        float u0 = rnd(rng);
        float u1 = rnd(rng);
        F3 h;
        if (!sampleVNDF(s.wo, s.g, u0, u1, h)) return false;
        float woh = dot(s.wo, h);
        if (woh <= 0.0f) return false;
        float F = fresnelDielectric(woh, etaFor(m, s.ff, s.l, s.spectral));
        float tp = transmissionProb(m, F);
        if (tp > 0.0f && rnd(rng) < tp) {
            float eta = refractionEtaFor(m, s.ff, s.l, s.spectral);
            F3 i = -s.wo;
            float ndi = dot(h, i);
            float k = 1.0f - eta * eta * (1.0f - ndi * ndi);
            F3 w = k < 0.0f ? f3(0.0f) : (i * eta - h * (eta * ndi + sqrtf(k)));
            if (dot(w, w) > 0.0f && w.z < 0.0f) {
                wi = w;
                isT = 1u;
                return true;
            }
        }
        wi = h * (2.0f * woh) - s.wo;
        isT = 0u;
        return wi.z > 0.0f;
    }
    float u1 = rnd(rng);
    float u2 = rnd(rng);
    float r = sqrtf(u1);
    float phi = 2.0f * kPi * u2;
    wi = F3{r * cosf(phi), r * sinf(phi), sqrtf(fmaxf(0.0f, 1.0f - u1))};
    return true;
}

}  // namespace mrt
