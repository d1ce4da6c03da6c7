// host/envmap.cpp - see envmap.h.
#include "host/envmap.h"
#include "host/mesh.h"
#include <cmath>
#include <cstdio>
#include <cstring>

namespace cwp {

namespace {

// RGBE -> float: mantissa * 2^(e - 136), no half-unit offset (the decoding Blender and the reference use,
// src/io/hdr_image.cpp:7-11)
void rgbeDecode(const unsigned char* e, float* out) {
    if (e[3] == 0) { out[0] = out[1] = out[2] = 0.f; return; }
    const float f = ldexpf(1.f, (int)e[3] - 136);
    out[0] = e[0] * f; out[1] = e[1] * f; out[2] = e[2] * f;
}

// float -> RGBE with the largest component's exponent; the mantissas are truncated, so rgbeDecode returns the value
// rounded down to 8 significant bits
void rgbeEncode(const float* c, unsigned char* e) {
    const float m = fmaxf(c[0], fmaxf(c[1], c[2]));
    if (!(m > 1e-32f)) { e[0] = e[1] = e[2] = e[3] = 0; return; }
    int ex = 0;
    frexpf(m, &ex);                           // m = f 2^ex, f in [0.5, 1)
    const float s = ldexpf(1.f, 8 - ex);      // 256 / 2^ex
    for (int k = 0; k < 3; k++) {
        const float v = fmaxf(c[k], 0.f) * s;
        e[k] = (unsigned char)(v >= 255.f ? 255 : (int)v);
    }
    e[3] = (unsigned char)(ex + 128);
}

struct FileCloser { FILE* f; ~FileCloser() { if (f) fclose(f); } };

}  // namespace

bool loadHDR(const std::string& path, int& w, int& h, std::vector<float>& rgb, std::string& err) {
    FILE* fp = fopen(path.c_str(), "rb");
    if (!fp) { err = "hdr: cannot open " + path; return false; }
    FileCloser guard{ fp };
    char line[512];
    if (!fgets(line, sizeof line, fp) || (strncmp(line, "#?RADIANCE", 10) != 0 && strncmp(line, "#?RGBE", 6) != 0)) {
        err = "hdr: " + path + " is not a Radiance file"; return false;
    }
    while (fgets(line, sizeof line, fp)) {
        if (line[0] == '\n' || line[0] == '\r') break;
        if (strncmp(line, "FORMAT=", 7) == 0 && strncmp(line + 7, "32-bit_rle_rgbe", 15) != 0) {
            err = "hdr: " + path + ": only FORMAT=32-bit_rle_rgbe is supported"; return false;
        }
    }
    if (!fgets(line, sizeof line, fp) || sscanf(line, "-Y %d +X %d", &h, &w) != 2 || w <= 0 || h <= 0 || w > 65536 || h > 65536) {
        err = "hdr: " + path + ": only the '-Y H +X W' orientation is supported"; return false;
    }
    rgb.assign((size_t)w * h * 3, 0.f);
    std::vector<unsigned char> scan((size_t)w * 4);
    for (int y = 0; y < h; y++) {
        unsigned char hd[4];
        if (fread(hd, 1, 4, fp) != 4) { err = "hdr: " + path + " truncated"; return false; }
        if (w >= 8 && w < 32768 && hd[0] == 2 && hd[1] == 2 && ((hd[2] << 8) | hd[3]) == w) {
            for (int c = 0; c < 4; c++) {   // new-style run length: each channel of the scanline separately
                int x = 0;
                while (x < w) {
                    int n = fgetc(fp);
                    if (n == EOF) { err = "hdr: " + path + " truncated"; return false; }
                    if (n > 128) {
                        n -= 128;
                        const int v = fgetc(fp);
                        if (v == EOF || x + n > w) { err = "hdr: " + path + " has a corrupt run"; return false; }
                        for (int k = 0; k < n; k++) scan[4 * (x++) + c] = (unsigned char)v;
                    } else {
                        if (n == 0 || x + n > w) { err = "hdr: " + path + " has a corrupt run"; return false; }
                        for (int k = 0; k < n; k++) {
                            const int v = fgetc(fp);
                            if (v == EOF) { err = "hdr: " + path + " truncated"; return false; }
                            scan[4 * (x++) + c] = (unsigned char)v;
                        }
                    }
                }
            }
        } else {   // flat scanline
            memcpy(scan.data(), hd, 4);
            if (w > 1 && fread(scan.data() + 4, 1, (size_t)(w - 1) * 4, fp) != (size_t)(w - 1) * 4) {
                err = "hdr: " + path + " truncated"; return false;
            }
        }
        for (int x = 0; x < w; x++) rgbeDecode(&scan[4 * x], &rgb[((size_t)y * w + x) * 3]);
    }
    return true;
}

bool writeHDR(const std::string& path, int w, int h, const std::vector<float>& rgb, std::string& err) {
    if (w < 1 || h < 1 || rgb.size() != (size_t)w * h * 3) { err = "hdr: image size mismatch"; return false; }
    FILE* fp = fopen(path.c_str(), "wb");
    if (!fp) { err = "hdr: cannot write " + path; return false; }
    FileCloser guard{ fp };
    fprintf(fp, "#?RADIANCE\nFORMAT=32-bit_rle_rgbe\n\n-Y %d +X %d\n", h, w);
    std::vector<unsigned char> scan((size_t)w * 4);
    for (int y = 0; y < h; y++) {
        for (int x = 0; x < w; x++) rgbeEncode(&rgb[((size_t)y * w + x) * 3], &scan[4 * x]);
        if (fwrite(scan.data(), 1, scan.size(), fp) != scan.size()) { err = "hdr: write failed"; return false; }
    }
    return true;
}

void buildEnvMap(const std::vector<float>& rgb, int w, int h, float strength, EnvMap& E) {
    E.w = w; E.h = h;
    E.texels.assign((size_t)w * h, v4(0.f, 0.f, 0.f, 0.f));
    E.marg.assign(h + 1, 0.f);
    E.cond.assign((size_t)h * (w + 1), 0.f);
    std::vector<double> rowSum(h, 0.0), rowCdf(w + 1), mc(h + 1, 0.0);
    for (int y = 0; y < h; y++) {
        const double st = std::sin(3.14159265358979323846 * (y + 0.5) / h);
        rowCdf[0] = 0.0;
        for (int x = 0; x < w; x++) {
            const float* c = &rgb[((size_t)y * w + x) * 3];
            const Vec3 L = v3(c[0], c[1], c[2]) * strength;
            E.texels[(size_t)y * w + x] = v4(L, 0.f);
            const double lum = 0.2126 * L.x + 0.7152 * L.y + 0.0722 * L.z;
            rowCdf[x + 1] = rowCdf[x] + (lum > 0 ? lum : 0.0) * st;
        }
        rowSum[y] = rowCdf[w];
        float* cr = &E.cond[(size_t)y * (w + 1)];
        for (int x = 0; x <= w; x++) cr[x] = rowSum[y] > 0 ? (float)(rowCdf[x] / rowSum[y]) : (float)x / (float)w;
        cr[w] = 1.f;
        mc[y + 1] = mc[y] + rowSum[y];
    }
    const double tot = mc[h];
    for (int y = 0; y <= h; y++) E.marg[y] = tot > 0 ? (float)(mc[y] / tot) : (float)y / (float)h;
    E.marg[h] = 1.f;
}

namespace {
// smooth value noise on the lattice of 3D integer points (trilinear, quintic fade)
double latticeValue(uint64_t seed, int x, int y, int z) {
    uint64_t k = seed ^ ((uint64_t)(uint32_t)x * 0x9E3779B97F4A7C15ull) ^ ((uint64_t)(uint32_t)y * 0xC2B2AE3D27D4EB4Full) ^
                 ((uint64_t)(uint32_t)z * 0x165667B19E3779F9ull);
    SplitMix r(k);
    return 2.0 * r.next01() - 1.0;
}
double valueNoise(uint64_t seed, double x, double y, double z) {
    const int ix = (int)std::floor(x), iy = (int)std::floor(y), iz = (int)std::floor(z);
    const double fx = x - ix, fy = y - iy, fz = z - iz;
    auto fade = [](double t) { return t * t * t * (t * (t * 6 - 15) + 10); };
    const double ux = fade(fx), uy = fade(fy), uz = fade(fz);
    double acc = 0;
    for (int c = 0; c < 8; c++) {
        const int dx = c & 1, dy = (c >> 1) & 1, dz = c >> 2;
        const double wgt = (dx ? ux : 1 - ux) * (dy ? uy : 1 - uy) * (dz ? uz : 1 - uz);
        acc += wgt * latticeValue(seed, ix + dx, iy + dy, iz + dz);
    }
    return acc;
}
}  // namespace

void makeLogNormalEnv(int w, int h, float mu, float sigma, uint64_t seed, std::vector<float>& rgb) {
    std::vector<double> z((size_t)w * h);
    double sum = 0, sum2 = 0;
    for (int y = 0; y < h; y++) {
        const double th = 3.14159265358979323846 * (y + 0.5) / h;
        for (int x = 0; x < w; x++) {
            const double ph = (0.5 - (x + 0.5) / w) * 2 * 3.14159265358979323846;
            const double dx = std::sin(th) * std::cos(ph), dy = std::sin(th) * std::sin(ph), dz = std::cos(th);
            double v = 0, amp = 1, f = 2.0;
            for (int o = 0; o < 6; o++) { v += amp * valueNoise(seed + (uint64_t)o, dx * f, dy * f, dz * f); amp *= 0.5; f *= 2.0; }
            z[(size_t)y * w + x] = v; sum += v; sum2 += v * v;
        }
    }
    const double n = (double)w * h, mean = sum / n, sd = std::sqrt(std::fmax(sum2 / n - mean * mean, 1e-30));
    rgb.assign((size_t)w * h * 3, 0.f);
    for (int y = 0; y < h; y++) {
        const double up = std::cos(3.14159265358979323846 * (y + 0.5) / h);   // +1 at the zenith
        const double t = 0.5 + 0.5 * up;
        const double tint[3] = { 0.85 + 0.3 * (1 - t), 0.95, 0.85 + 0.3 * t };   // warm below, blue above
        const double tl = 0.2126 * tint[0] + 0.7152 * tint[1] + 0.0722 * tint[2];
        for (int x = 0; x < w; x++) {
            const double L = std::exp(mu + sigma * (z[(size_t)y * w + x] - mean) / sd);
            for (int k = 0; k < 3; k++) rgb[((size_t)y * w + x) * 3 + k] = (float)(L * tint[k] / tl);
        }
    }
}

}  // namespace cwp
