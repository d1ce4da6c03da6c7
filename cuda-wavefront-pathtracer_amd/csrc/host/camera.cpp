// host/camera.cpp - camera construction and the pixel-filter table.
// Follows src/scene/camera.cpp:6-23 (look-at frame, focal length from a horizontal field of view, principal point at
// the image centre) and src/core/sampling.h:32-43 (Blackman-Harris inverse CDF built in double precision from a
// 4096-cell numerical CDF).
#include "host/camera.h"
#include <cmath>
#include <vector>

namespace cwp {

void identityC2W(float c2w[16]) { for (int j = 0; j < 16; j++) c2w[j] = (j % 5 == 0) ? 1.f : 0.f; }

void lookAtC2W(Vec3 pos, Vec3 tgt, Vec3 up, float c2w[16]) {
    identityC2W(c2w);
    Vec3 f = normalize(tgt - pos), r = normalize(cross(f, up)), u = cross(r, f), back = -f;
    const Vec3 cols[3] = { r, u, back };
    for (int row = 0; row < 3; row++)
        for (int col = 0; col < 3; col++) c2w[row * 4 + col] = comp(cols[col], row);
    c2w[3] = pos.x; c2w[7] = pos.y; c2w[11] = pos.z;
}

float focalFromHfov(int width, float hfovDeg) { return (width * 0.5f) / tanf(hfovDeg * 0.5f * 3.14159265f / 180.f); }

Camera makeCamera(const float c2w[16], float fx, float fy, int w, int h) {
    Camera c;
    c.pos = v3(c2w[3], c2w[7], c2w[11]);
    c.ax = v3(c2w[0], c2w[4], c2w[8]); c.ay = v3(c2w[1], c2w[5], c2w[9]); c.az = v3(c2w[2], c2w[6], c2w[10]);
    c.fx = fx; c.fy = fy; c.px = w * 0.5f; c.py = h * 0.5f; c.w = w; c.h = h; c.lensR = 0.f; c.focusD = 1.f;
    return c;
}

void blackmanHarrisTable(float* tab, int n, double width) {
    const int cells = 4096;
    std::vector<double> cdf(cells + 1, 0.0);
    for (int i = 0; i < cells; i++) {
        double t = (i + 0.5) / cells;
        double f = 0.35875 - 0.48829 * cos(2 * M_PI * t) + 0.14128 * cos(4 * M_PI * t) - 0.01168 * cos(6 * M_PI * t);
        cdf[i + 1] = cdf[i] + fmax(f, 0.0);
    }
    for (int i = 0; i <= cells; i++) cdf[i] /= cdf[cells];
    for (int k = 0; k <= n; k++) {
        double u = (double)k / n;
        int j = 0;
        while (j < cells - 1 && cdf[j + 1] < u) j++;
        double t = cdf[j + 1] > cdf[j] ? (u - cdf[j]) / (cdf[j + 1] - cdf[j]) : 0.0;
        tab[k] = (float)(((j + t) / cells - 0.5) * width);
    }
}

}  // namespace cwp
