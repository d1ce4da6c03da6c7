// host/camera.h - pinhole / thin-lens camera, lighting, and the per-frame parameter block of the legacy integrator.
#pragma once
#include "core/sampling.h"
#include <string>

namespace cwp {

// OpenGL camera convention (looks down -Z, +Y up) in a z-up world; lensR = 0 is the pinhole camera
// (reference: src/scene/camera.h:7).
struct Camera {
    Vec3 pos, ax, ay, az;
    float fx, fy, px, py;
    int w, h;
    float lensR, focusD;
};

// Directional sun with a finite angular diameter plus a constant sky (reference: src/scene/lighting.h:6-14).
struct Lighting {
    Vec3 sunTo = v3(0.3f, -0.5f, 0.8f);   // direction toward the sun (normalised by the setup)
    Vec3 sunE = v3(3.f, 3.f, 3.f);        // irradiance on a surface facing the sun
    float sunAngleDeg = 0.526f;           // angular diameter
    Vec3 sky = v3(0.3f, 0.3f, 0.3f);      // constant sky radiance (no environment map)
    std::string envMap;                   // equirectangular Radiance .hdr map (replaces the sky; full integrator)
    float envStrength = 1.f;              // radiance multiplier of the map
    float envRotDeg = 0.f;                // rotation about the world z axis
};

// Parameters of one frame of the legacy integrator, passed by value as a kernel argument (reference: the __constant__
// Params block, src/render/params.h:8-16).
struct FrameParams {
    Camera cam;
    Vec3 sunTo, sunE, sky;
    float sunCos;            // cos of the sun's angular radius
    int maxB;                // paths stop after bounce maxB (camera hit = depth 0)
    uint32_t seed;
    float eps, nearT;        // ray-origin offset (world units), camera near clip
    int hasN, filterBH;      // shading normals present; Blackman-Harris pixel filter (else 1 px box)
    float ftab[kFilterTable + 1];
};

void identityC2W(float c2w[16]);
void lookAtC2W(Vec3 pos, Vec3 tgt, Vec3 up, float c2w[16]);
float focalFromHfov(int width, float hfovDeg);
Camera makeCamera(const float c2w[16], float fx, float fy, int w, int h);
// inverse CDF of the 1D Blackman-Harris window of total width `width` pixels, tabulated at n + 1 points
void blackmanHarrisTable(float* tab, int n, double width);

}  // namespace cwp
