// core/texture.h - texture maps of the full integrator, shared by host and device: per-material bindings, image
// records and the lookup.
//
// Semantics follow the reference's glTF metallic-roughness model (src/scene/texture.h:15-25,
// src/render/wavefront.cu:199-204,252-271): maps multiply the material's factors (base colour and emission by the
// map's rgb, metalness and roughness by one channel of theirs, glTF packing metalness in B and roughness in G), and a
// tangent-space normal map n = T (2r - 1) s + B (2g - 1) s + N (2b - 1) with strength s. Texture coordinates have
// their origin at the bottom left (v up); images are stored top row first, so a lookup reads row (1 - v) h.
//
// The lookup is this build's own, evaluated identically on the host and the device (the reference uses the CUDA
// texture unit): normalised coordinates wrap (t - floor t); nearest picks texel floor(t * size); linear blends the 2x2
// texels around t * size - 0.5 with weights from its fraction, each texel decoded first (colour maps from sRGB through
// a 256-entry table, the others as byte / 255).
#pragma once
#include "core/vecmath.h"
#include <cstdint>

namespace cwp {

// Texture bindings of one material (32 bytes): image ids (-1: none) and the normal map's strength.
struct MaterialMaps {
    int baseColor, emission, normal, metallic, roughness;
    int metallicChannel, roughnessChannel;   // 0..3 = r, g, b, a
    float normalScale;
};
static_assert(sizeof(MaterialMaps) == 32, "MaterialMaps is 32 bytes (two float4 loads)");

CWP_HD MaterialMaps noMaps() {
    MaterialMaps m;
    m.baseColor = m.emission = m.normal = m.metallic = m.roughness = -1;
    m.metallicChannel = 2; m.roughnessChannel = 1; m.normalScale = 1.f;
    return m;
}
CWP_HD bool hasMaps(const MaterialMaps& m) {
    return m.baseColor >= 0 || m.emission >= 0 || m.normal >= 0 || m.metallic >= 0 || m.roughness >= 0;
}

// One image in the texel array: first texel, size, decode (1: colour data stored as sRGB).
struct TexImage {
    uint32_t first;
    int w, h, srgb;
};

CWP_HD float texelChannel(uint32_t px, int c, int srgb, const float* srgbTable) {
    const uint32_t b = (px >> (8 * c)) & 255u;
    return (srgb && c < 3) ? srgbTable[b] : (float)b * (1.f / 255.f);
}
CWP_HD Vec4 texelValue(uint32_t px, int srgb, const float* srgbTable) {
    return v4(texelChannel(px, 0, srgb, srgbTable), texelChannel(px, 1, srgb, srgbTable),
              texelChannel(px, 2, srgb, srgbTable), texelChannel(px, 3, srgb, srgbTable));
}
CWP_HD float lerpf(float a, float b, float t) { return fmaf(t, b - a, a); }
CWP_HD Vec4 lerp4(Vec4 a, Vec4 b, float t) {
    return v4(lerpf(a.x, b.x, t), lerpf(a.y, b.y, t), lerpf(a.z, b.z, t), lerpf(a.w, b.w, t));
}

// RGBA of image T at texture coordinates (u, v); texels: RGBA8 words (r in the low byte), rows top first.
CWP_HD Vec4 sampleTexture(const TexImage& T, const uint32_t* texels, const float* srgbTable, float u, float v,
                          int nearest) {
    const float vv = 1.f - v;
    float tu = u - floorf(u), tv = vv - floorf(vv);   // in [0, 1] (1 for tiny negative inputs); NaN / inf -> 0
    if (!(tu >= 0.f && tu <= 1.f)) tu = 0.f;
    if (!(tv >= 0.f && tv <= 1.f)) tv = 0.f;
    const uint32_t* px = texels + T.first;
    if (nearest) {
        const int xi = (int)(tu * (float)T.w), yi = (int)(tv * (float)T.h);
        const int x = xi < T.w - 1 ? xi : T.w - 1, y = yi < T.h - 1 ? yi : T.h - 1;
        return texelValue(px[(size_t)y * T.w + x], T.srgb, srgbTable);
    }
    const float x = tu * (float)T.w - 0.5f, y = tv * (float)T.h - 0.5f;
    const float fx0 = floorf(x), fy0 = floorf(y);
    const float ax = x - fx0, ay = y - fy0;
    int x0 = (int)fx0, y0 = (int)fy0;
    if (x0 < 0) x0 += T.w;   // x >= -0.5: only the left border wraps
    if (y0 < 0) y0 += T.h;
    const int x1 = x0 + 1 < T.w ? x0 + 1 : 0, y1 = y0 + 1 < T.h ? y0 + 1 : 0;
    const Vec4 c00 = texelValue(px[(size_t)y0 * T.w + x0], T.srgb, srgbTable);
    const Vec4 c10 = texelValue(px[(size_t)y0 * T.w + x1], T.srgb, srgbTable);
    const Vec4 c01 = texelValue(px[(size_t)y1 * T.w + x0], T.srgb, srgbTable);
    const Vec4 c11 = texelValue(px[(size_t)y1 * T.w + x1], T.srgb, srgbTable);
    return lerp4(lerp4(c00, c10, ax), lerp4(c01, c11, ax), ay);
}

CWP_HD float channelOf(Vec4 c, int ch) { return ch == 0 ? c.x : (ch == 1 ? c.y : (ch == 2 ? c.z : c.w)); }

}  // namespace cwp
