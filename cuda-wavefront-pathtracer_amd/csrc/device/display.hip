// device/display.hip - the viewer's display transform on the GPU: accumulated radiance sums -> packed RGBA8
// (reference: k_display / launchDisplay, src/app/display.cu:21-33). One thread per pixel, 256-thread workgroups; the
// 255 step positions of the 8-bit sRGB encode travel with the launch (a 1 KiB kernel argument) and are staged in LDS,
// so a channel's encode is 8 LDS compares instead of a power function. The kernel reads 16 and writes 4 bytes per
// pixel.
#include "core/display.h"
#include "device/display.h"
#include "host/io.h"
#include <hip/hip_runtime.h>
#include <cmath>
#include <mutex>

namespace cwp {
namespace dev {

struct SrgbStepArg { float t[kSrgbSteps]; };

constexpr int kDisplayBlock = 256;
static_assert(kDisplayBlock == kSrgbSteps, "one step position per thread");

__global__ void __launch_bounds__(kDisplayBlock) k_display(const Vec4* __restrict__ acc, uint32_t* __restrict__ out,
                                                           int n, float spp, float scale, int mode, SrgbStepArg steps) {
    __shared__ float t[kSrgbSteps];
    t[threadIdx.x] = steps.t[threadIdx.x];
    __syncthreads();
    const int i = blockIdx.x * kDisplayBlock + (int)threadIdx.x;
    if (i < n) out[i] = displayPixel(t, acc[i], spp, scale, mode);
}

}  // namespace dev

static const dev::SrgbStepArg& stepTable() {
    static dev::SrgbStepArg T;
    static std::once_flag once;
    std::call_once(once, [] { srgbSteps(T.t); });
    return T;
}

bool displayLaunch(const Vec4* accum, uint32_t* out, int n, int spp, float exposureEV, int mode, void* stream,
                   std::string& err) {
    if (n < 0 || (n > 0 && (!accum || !out))) { err = "display: bad buffers"; return false; }
    if (mode < kDisplayClip || mode > kDisplayACES) { err = "display: mode must be 0 (clip), 1 (reinhard) or 2 (aces)"; return false; }
    if (!std::isfinite(exposureEV)) { err = "display: exposure must be finite"; return false; }
    if (spp < 0) { err = "display: spp must be >= 0"; return false; }
    if (n == 0) return true;
    const float scale = exposureEV == 0.f ? 1.f : exp2f(exposureEV);
    const int blocks = (n + dev::kDisplayBlock - 1) / dev::kDisplayBlock;
    dev::k_display<<<blocks, dev::kDisplayBlock, 0, (hipStream_t)stream>>>(accum, out, n, (float)(spp > 0 ? spp : 1),
                                                                           scale, mode, stepTable());
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) { err = std::string("display: ") + hipGetErrorString(e); return false; }
    return true;
}

bool displaySync(const Vec4* accum, uint32_t* out, int n, int spp, float exposureEV, int mode, std::string& err) {
    if (!displayLaunch(accum, out, n, spp, exposureEV, mode, nullptr, err)) return false;
    const hipError_t e = hipStreamSynchronize(nullptr);
    if (e != hipSuccess) { err = std::string("display: ") + hipGetErrorString(e); return false; }
    return true;
}

}  // namespace cwp
