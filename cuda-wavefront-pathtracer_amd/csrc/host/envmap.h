// host/envmap.h - equirectangular environment maps: Radiance .hdr files, the generated log-normal map, and the
// piecewise-constant importance-sampling distribution of the full integrator.
//
// Orientation (the reference's, src/scene/environment.h:1-5): row 0 is +z (up); the centre column faces +x and u grows
// toward -y: u = 0.5 - atan2(y, x) / 2 pi, v = acos(z) / pi. Distribution: texel weight = luminance * sin(theta) of
// the texel's row centre, a marginal CDF over rows (h + 1 entries) and a conditional CDF per row (w + 1 entries each),
// accumulated in double and rounded once (reference: buildEnvMap, src/scene/environment.cpp:7-37).
#pragma once
#include "core/vecmath.h"
#include <cstdint>
#include <string>
#include <vector>

namespace cwp {

struct EnvMap {
    int w = 0, h = 0;
    std::vector<Vec4> texels;   // radiance * strength, row-major, row 0 = top
    std::vector<float> marg;    // h + 1
    std::vector<float> cond;    // h rows of w + 1
};

// Radiance RGBE (.hdr, "#?RADIANCE" / "#?RGBE", FORMAT=32-bit_rle_rgbe, -Y h +X w): flat and new-style run-length
// scanlines. rgb: w * h * 3 floats, row 0 = top.
bool loadHDR(const std::string& path, int& w, int& h, std::vector<float>& rgb, std::string& err);
// writes flat (uncompressed) RGBE scanlines
bool writeHDR(const std::string& path, int w, int h, const std::vector<float>& rgb, std::string& err);

// texels = rgb * strength and the sampling distribution
void buildEnvMap(const std::vector<float>& rgb, int w, int h, float strength, EnvMap& E);

// Generated environment: per-texel luminance log-normal with ln L = mu + sigma z, z a zero-mean, unit-variance fBm
// over the sphere (value noise on the direction, 6 octaves), tinted toward blue above the horizon and warm below.
void makeLogNormalEnv(int w, int h, float mu, float sigma, uint64_t seed, std::vector<float>& rgb);

}  // namespace cwp
