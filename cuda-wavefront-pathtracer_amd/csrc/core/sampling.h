// core/sampling.h - stateless hash RNG and the sample warps of the legacy integrator (host and device).
//
// The RNG must produce the reference's stream so that a path keyed by (pixel, sample, depth, dimension, seed) takes the
// same random decisions as the reference does: the hash constants and the key mixing are those of
// src/core/rng.h:6-12. The warps restate src/core/sampling.h:13-58 (Duff et al. 2017 orthonormal basis, cosine
// hemisphere, uniform cone, piecewise-linear inverse-CDF table, concentric disk, uniform sphere).
#pragma once
#include "core/vecmath.h"

namespace cwp {

constexpr float kTwoPi = 6.2831853f;
constexpr float kInvPi = 0.31830989f;
constexpr int kFilterTable = 64;   // inverse-CDF table of the Blackman-Harris pixel filter has kFilterTable + 1 points

CWP_HD uint32_t hashU32(uint32_t x) {
    x ^= x >> 16; x *= 0x7feb352dU; x ^= x >> 15; x *= 0x846ca68bU; x ^= x >> 16; return x;
}

// uniform float in [0, 1) with 24 bits; dimension index = depth * 8 + dim
CWP_HD float rnd(uint32_t pix, uint32_t smp, uint32_t depth, uint32_t dim, uint32_t seed) {
    uint32_t inner = hashU32((depth * 8u + dim) * 0xC2B2AE3DU ^ seed);
    uint32_t mid = hashU32(smp * 0x85EBCA77U ^ inner);
    uint32_t h = hashU32(pix * 0x9E3779B1U ^ mid);
    return (float)(h >> 8) * (1.0f / 16777216.0f);
}

// orthonormal basis (t, b, n) around a unit normal n
CWP_HD void basis(Vec3 n, Vec3& t, Vec3& b) {
    float s = n.z >= 0.f ? 1.f : -1.f;
    float a = -1.f / (s + n.z);
    float c = n.x * n.y * a;
    t = v3(1.f + s * n.x * n.x * a, s * c, -s * n.x);
    b = v3(c, s + n.y * n.y * a, -n.y);
}

// cosine-weighted direction around n (pdf cos/pi)
CWP_HD Vec3 cosineHemisphere(Vec3 n, float u0, float u1) {
    float r = sqrtf(u0), ph = kTwoPi * u1;
    Vec3 t, b; basis(n, t, b);
    Vec3 a = t * (r * cosf(ph));
    Vec3 c = b * (r * sinf(ph));
    Vec3 z = n * sqrtf(fmaxf(0.f, 1.f - u0));
    return normalize(a + c + z);
}

// uniform direction in the cone of half-angle acos(cosMax) around the unit axis
CWP_HD Vec3 uniformCone(Vec3 axis, float cosMax, float u0, float u1) {
    float cz = 1.f - u0 * (1.f - cosMax);
    float sz = sqrtf(fmaxf(0.f, 1.f - cz * cz));
    float ph = kTwoPi * u1;
    Vec3 t, b; basis(axis, t, b);
    Vec3 a = axis * cz;
    Vec3 c = t * (sz * cosf(ph));
    Vec3 e = b * (sz * sinf(ph));
    return normalize(a + c + e);
}

// piecewise-linear lookup into an inverse CDF of n + 1 entries, u in [0, 1)
CWP_HD float tableLookup(const float* tab, int n, float u) {
    float f = u * (float)n;
    int k = (int)f; k = k < n - 1 ? k : n - 1;
    float t = f - (float)k;
    return tab[k] * (1.f - t) + tab[k + 1] * t;
}

// uniform direction on the unit sphere
CWP_HD Vec3 uniformSphere(float u0, float u1) {
    const float z = 1.f - 2.f * u0;
    const float r = sqrtf(fmaxf(0.f, 1.f - z * z));
    const float ph = kTwoPi * u1;
    return v3(r * cosf(ph), r * sinf(ph), z);
}

// uniform point on the unit disk (Shirley-Chiu concentric map)
CWP_HD void concentricDisk(float u0, float u1, float& x, float& y) {
    float a = 2.f * u0 - 1.f, b = 2.f * u1 - 1.f;
    if (a == 0.f && b == 0.f) { x = 0.f; y = 0.f; return; }
    float r, ph;
    if (fabsf(a) > fabsf(b)) { r = a; ph = 0.78539816f * (b / a); }
    else { r = b; ph = 1.57079633f - 0.78539816f * (a / b); }
    x = r * cosf(ph); y = r * sinf(ph);
}

}  // namespace cwp
