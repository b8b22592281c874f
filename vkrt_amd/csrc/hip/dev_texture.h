// Material textures on the device: bilinear filtering of mip level 0 with
// per-axis repeat / clamp-to-edge / mirrored-repeat addressing (the nine
// sampler variants of src/core/scene/textures.c:141-178: linear min/mag
// filter, maxLod 0), the UV transform and texture-coordinate set selection of
// src/shaders/material/textures.slang:8-121, and the texture stages
// applySurfaceTextures / applyNormalTexture (textures.slang:123-155).
//
// The pool holds every referenced texture as linear RGBA32F (sRGB bytes are
// decoded per texel at upload, before filtering, as an sRGB image view does).
// The filter weights are full-precision floats rather than a texture unit's
// fixed-point sub-texel weights.
#pragma once

#include "dev_common.h"

namespace mrt {

// Texel index along one axis for a glTF wrap code (anything else repeats,
// textures.slang:16-26).
DEV int wrapTexel(int i, int n, uint32_t mode) {
    if (mode == VKRT_TEXTURE_WRAP_CLAMP_TO_EDGE) return min(max(i, 0), n - 1);
    if (mode == VKRT_TEXTURE_WRAP_MIRRORED_REPEAT) {
        const int p = 2 * n;
        int t = i % p;
        if (t < 0) t += p;
        return t < n ? t : p - 1 - t;
    }
    int t = i % n;
    return t < 0 ? t + n : t;
}

// Texel-space coordinate, bounded so that the integer conversion is defined
// for every input (a NaN coordinate becomes the lower bound).
DEV float texelCoord(float uv, uint32_t size) {
    const float x = uv * (float)size - 0.5f;
    return fminf(fmaxf(x, -16777216.0f), 16777216.0f);
}

DEV F4 ldTexel(const SceneView& s, uint32_t base, uint32_t w, int x, int y) {
    const float4 t = s.texels[base + (uint32_t)y * w + (uint32_t)x];
    return F4{t.x, t.y, t.z, t.w};
}

DEV F4 sampleTexture(const SceneView& s, uint32_t tex, float u, float v, uint32_t packedWrap) {
    const uint4 info = s.texInfo[tex];
    const int w = (int)info.y, h = (int)info.z;
    const float x = texelCoord(u, info.y), y = texelCoord(v, info.z);
    const float fx = floorf(x), fy = floorf(y);
    const float ax = x - fx, ay = y - fy;
    const int i0 = (int)fx, j0 = (int)fy;
    const uint32_t wu = packedWrap & 0xffffu, wv = (packedWrap >> 16) & 0xffffu;
    const int xa = wrapTexel(i0, w, wu), xb = wrapTexel(i0 + 1, w, wu);
    const int ya = wrapTexel(j0, h, wv), yb = wrapTexel(j0 + 1, h, wv);
    const F4 t00 = ldTexel(s, info.x, info.y, xa, ya), t10 = ldTexel(s, info.x, info.y, xb, ya);
    const F4 t01 = ldTexel(s, info.x, info.y, xa, yb), t11 = ldTexel(s, info.x, info.y, xb, yb);
    const F4 top = t00 * (1.0f - ax) + t10 * ax;
    const F4 bottom = t01 * (1.0f - ax) + t11 * ax;
    return top * (1.0f - ay) + bottom * ay;
}

// Interpolated texture coordinates of a hit (interpolation.slang:46-60).
struct TexCoords {
    float u0, v0, u1, v1;
};

DEV TexCoords hitTexCoords(const SceneView& s, uint32_t tri, uint32_t mesh, float u, float v) {
    const MeshInfo& m = s.meshes[mesh];
    const uint32_t tb = m.indexBase + (tri - s.meshTriBase[mesh]) * 3u;
    const ShaderVertex& a = s.vertices[s.indices[tb + 0] + m.vertexBase];
    const ShaderVertex& b = s.vertices[s.indices[tb + 1] + m.vertexBase];
    const ShaderVertex& c = s.vertices[s.indices[tb + 2] + m.vertexBase];
    const float w = 1.0f - u - v;
    TexCoords r;
    r.u0 = a.texcoord0[0] * w + b.texcoord0[0] * u + c.texcoord0[0] * v;
    r.v0 = a.texcoord0[1] * w + b.texcoord0[1] * u + c.texcoord0[1] * v;
    r.u1 = a.texcoord1[0] * w + b.texcoord1[0] * u + c.texcoord1[0] * v;
    r.v1 = a.texcoord1[1] * w + b.texcoord1[1] * u + c.texcoord1[1] * v;
    return r;
}

// sampleMaterialTexture (textures.slang:58-74): the slot's coordinate set,
// scale / rotation / offset, then the filtered sample; an empty slot gives the
// fallback value.
DEV F4 sampleMaterialSlot(const SceneView& s, const Material& m, uint32_t slot, const TexCoords& tc, F4 fallback) {
    const uint32_t tex = slot == 0u ? m.baseColorTextureIndex
                                    : (slot == 1u ? m.metallicRoughnessTextureIndex
                                                  : (slot == 2u ? m.normalTextureIndex : m.emissiveTextureIndex));
    if (tex == VKRT_INVALID_INDEX) return fallback;
    const uint32_t wrap = slot == 0u ? m.baseColorTextureWrap
                                     : (slot == 1u ? m.metallicRoughnessTextureWrap
                                                   : (slot == 2u ? m.normalTextureWrap : m.emissiveTextureWrap));
    const float* xf = slot == 0u ? m.baseColorTextureTransform
                                 : (slot == 1u ? m.metallicRoughnessTextureTransform
                                               : (slot == 2u ? m.normalTextureTransform : m.emissiveTextureTransform));
    const bool set1 = ((m.textureTexcoordSets >> (slot * 8u)) & 0xffu) == 1u;
    const float su = (set1 ? tc.u1 : tc.u0) * xf[0], sv = (set1 ? tc.v1 : tc.v0) * xf[1];
    const float rot = m.textureRotations[slot];
    const float sn = sinf(rot), cs = cosf(rot);
    return sampleTexture(s, tex, cs * su - sn * sv + xf[2], sn * su + cs * sv + xf[3], wrap);
}

DEV bool materialHasTextures(const Material& m) {
    return (m.baseColorTextureIndex & m.metallicRoughnessTextureIndex & m.normalTextureIndex &
            m.emissiveTextureIndex) != VKRT_INVALID_INDEX;
}

}  // namespace mrt
