// device/display.h - launcher of the display transform kernel (device/display.hip; semantics in core/display.h).
#pragma once
#include "core/vecmath.h"
#include <cstdint>
#include <string>

namespace cwp {

// Queues the transform of n accumulated radiance sums of spp samples (spp 0 is read as 1) into n packed RGBA8 pixels
// (r in the low byte, a = 255) on `stream` (a hipStream_t; null: the default stream). accum and out are device
// pointers. Mode: 0 clip, 1 Reinhard, 2 ACES; the exposure in stops scales the sums by 2^exposureEV.
bool displayLaunch(const Vec4* accum, uint32_t* out, int n, int spp, float exposureEV, int mode, void* stream,
                   std::string& err);
// The same on the default stream, returning when it is done.
bool displaySync(const Vec4* accum, uint32_t* out, int n, int spp, float exposureEV, int mode, std::string& err);

}  // namespace cwp
