// host/camera.cpp - camera construction and the pixel-filter table.
// Follows src/scene/camera.cpp:6-23 (look-at frame, focal length from a horizontal field of view, principal point at
// the image centre) and src/core/sampling.h:32-43 (Blackman-Harris inverse CDF built in double precision from a
// 4096-cell numerical CDF).
#include "host/camera.h"
#include <algorithm>
#include <cmath>
#include <vector>

namespace cwp {

void identityC2W(float c2w[16]) {
    std::fill(c2w, c2w + 16, 0.f);
    for (int k = 0; k < 4; k++) c2w[4 * k + k] = 1.f;
}

// camera-to-world rows: x = right, y = up, z = backward (the camera looks down -z), translation = position
void lookAtC2W(Vec3 pos, Vec3 tgt, Vec3 up, float c2w[16]) {
    const Vec3 fwd = normalize(tgt - pos);
    const Vec3 right = normalize(cross(fwd, up));
    const Vec3 upv = cross(right, fwd);
    const Vec3 back = -fwd;
    identityC2W(c2w);
    for (int r = 0; r < 3; r++) {
        float* row = c2w + 4 * r;
        row[0] = comp(right, r); row[1] = comp(upv, r); row[2] = comp(back, r); row[3] = comp(pos, r);
    }
}

// pinhole focal length (pixels) whose horizontal field of view over `width` pixels is hfovDeg
float focalFromHfov(int width, float hfovDeg) { return (width * 0.5f) / tanf(hfovDeg * 0.5f * 3.14159265f / 180.f); }

Camera makeCamera(const float c2w[16], float fx, float fy, int w, int h) {
    auto column = [&](int c) { return v3(c2w[c], c2w[4 + c], c2w[8 + c]); };
    Camera c;
    c.ax = column(0); c.ay = column(1); c.az = column(2); c.pos = column(3);
    c.fx = fx; c.fy = fy;
    c.px = w * 0.5f; c.py = h * 0.5f;   // principal point at the image centre
    c.w = w; c.h = h;
    c.lensR = 0.f; c.focusD = 1.f;      // pinhole until a lens is set
    return c;
}

// Inverse CDF (n + 1 points over [-width / 2, width / 2]) of the Blackman-Harris window: the window integrated over
// 4096 cells in double precision, then inverted by one sweep over the cells (the cell holding u only moves right as u
// grows), linear inside the cell.
void blackmanHarrisTable(float* tab, int n, double width) {
    const int cells = 4096;
    std::vector<double> cdf(cells + 1, 0.0);
    for (int i = 0; i < cells; i++) {
        const double t = (i + 0.5) / cells;   // cell centre
        const double w = 0.35875 - 0.48829 * cos(2 * M_PI * t) + 0.14128 * cos(4 * M_PI * t) - 0.01168 * cos(6 * M_PI * t);
        cdf[i + 1] = cdf[i] + fmax(w, 0.0);
    }
    const double total = cdf[cells];
    for (double& v : cdf) v /= total;
    int j = 0;
    for (int k = 0; k <= n; k++) {
        const double u = (double)k / n;
        while (j < cells - 1 && cdf[j + 1] < u) j++;
        const double lo = cdf[j], hi = cdf[j + 1];
        const double t = hi > lo ? (u - lo) / (hi - lo) : 0.0;
        tab[k] = (float)(((j + t) / cells - 0.5) * width);
    }
}

}  // namespace cwp
