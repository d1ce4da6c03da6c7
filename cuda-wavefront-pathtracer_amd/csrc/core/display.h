// core/display.h - the display transform of the progressive viewer: accumulated radiance sums -> 8-bit RGBA, shared
// by host and device.
//
// Semantics follow the reference's viewer display (src/app/display.cu:9-33): each channel is sum / spp, times
// 2^exposure when the exposure is not 0, then (modes 1 and 2, on max(x, 0)) a tone curve, Reinhard x / (1 + x) or the
// ACES filmic fit of Narkowicz (2015) with a 0.6 pre-scale; last the PNG writer's 8-bit encode: clip to [0, 1], the
// sRGB OETF, x * 255 + 0.5 truncated. Mode 0 at exposure 0 therefore gives the batch renderer's PNG bytes.
//
// The device does not evaluate the OETF's power function: the 8-bit encode is a step function of x, so it is applied
// as a search over its 255 step positions (the smallest float that encodes to each byte value), which the host
// computes once from its own encoder. The bytes then equal the host encoder's for every input.
#pragma once
#include "core/vecmath.h"
#include <cstdint>

namespace cwp {

enum DisplayMode { kDisplayClip = 0, kDisplayReinhard = 1, kDisplayACES = 2 };

// step positions of the 8-bit encode: t[k] (k = 1..255) is the smallest float that encodes to k or more; t[0] = -inf
constexpr int kSrgbSteps = 256;

CWP_HD float displayCurve(float x, int mode) {
    if (mode == kDisplayReinhard) return x / (1.f + x);
    if (mode == kDisplayACES) {
        const float s = x * 0.6f;
        return (s * (2.51f * s + 0.03f)) / (s * (2.43f * s + 0.59f) + 0.14f);
    }
    return x;
}

// the 8-bit encode of x through the step table: the largest k with x >= t[k] (NaN and x <= 0 give 0, x >= 1 gives 255)
CWP_HD uint32_t encodeBySteps(const float* t, float x) {
    x = fminf(fmaxf(x, 0.f), 1.f);
    uint32_t k = 0;
#pragma unroll
    for (uint32_t s = 128; s > 0; s >>= 1)
        if (x >= t[k + s]) k += s;
    return k;
}

// One pixel: the sums (r, g, b) of spp samples -> packed RGBA8 (a = 255; r in the low byte)
CWP_HD uint32_t displayPixel(const float* t, Vec4 a, float spp, float scale, int mode) {
    float c[3] = { a.x / spp, a.y / spp, a.z / spp };
    uint32_t out = 0xFF000000u;
#pragma unroll
    for (int k = 0; k < 3; k++) {
        float x = c[k];
        if (scale != 1.f) x = x * scale;
        if (mode != kDisplayClip) x = displayCurve(fmaxf(x, 0.f), mode);
        out |= encodeBySteps(t, x) << (8 * k);
    }
    return out;
}

}  // namespace cwp
