// host/mesh.cpp - mesh helpers and procedural scenes.
#include "host/mesh.h"
#include "host/camera.h"
#include <algorithm>
#include <cmath>
#include <map>

namespace cwp {

AABB triBox(const Mesh& m, size_t t) {
    AABB b; b.grow(m.v[m.f[t].a]); b.grow(m.v[m.f[t].b]); b.grow(m.v[m.f[t].c]); return b;
}
AABB meshBox(const Mesh& m) { AABB b; for (const Vec3& p : m.v) b.grow(p); return b; }

void addBox(Mesh& m, Vec3 lo, Vec3 hi, int mat) {
    int base = (int)m.v.size();
    for (int k = 0; k < 8; k++) m.v.push_back(v3(k & 1 ? hi.x : lo.x, k & 2 ? hi.y : lo.y, k & 4 ? hi.z : lo.z));
    // faces as corner-index quads (a, b, c, d) counter-clockwise seen from outside
    static const int q[6][4] = { { 0, 2, 3, 1 }, { 4, 5, 7, 6 }, { 0, 1, 5, 4 }, { 2, 6, 7, 3 }, { 0, 4, 6, 2 }, { 1, 3, 7, 5 } };
    for (int fc = 0; fc < 6; fc++) {
        int a = base + q[fc][0], b = base + q[fc][1], c = base + q[fc][2], d = base + q[fc][3];
        m.f.push_back({ a, b, c }); m.fm.push_back(mat);
        m.f.push_back({ a, c, d }); m.fm.push_back(mat);
    }
}

void makeBoxCloud(Mesh& m, int nBoxes, uint64_t seed, float albedo) {
    m = Mesh();
    m.kd.push_back(v3(albedo, albedo, albedo)); m.matName.push_back("diffuse");
    SplitMix r(seed);
    for (int i = 0; i < nBoxes; i++) {
        Vec3 c = v3(r.next01(), r.next01(), r.next01());
        Vec3 e = v3(r.uniform(0.01f, 0.08f), r.uniform(0.01f, 0.08f), r.uniform(0.01f, 0.08f));
        addBox(m, c - e * 0.5f, c + e * 0.5f, 0);
    }
    int g = (int)m.v.size();
    m.v.push_back(v3(-1.f, -1.f, 0.f)); m.v.push_back(v3(2.f, -1.f, 0.f)); m.v.push_back(v3(2.f, 2.f, 0.f)); m.v.push_back(v3(-1.f, 2.f, 0.f));
    m.f.push_back({ g, g + 1, g + 2 }); m.fm.push_back(0);
    m.f.push_back({ g, g + 2, g + 3 }); m.fm.push_back(0);
}

void boxCloudCamera(float c2w[16]) { lookAtC2W(v3(0.5f, -1.25f, 1.35f), v3(0.5f, 0.5f, 0.35f), v3(0.f, 0.f, 1.f), c2w); }

// value noise on the integer lattice: hashed corner values in [0,1), smoothstep interpolation (double precision)
static double latticeValue(int64_t x, int64_t y, uint64_t seed) {
    uint64_t h = (uint64_t)x * 0x8DA6B343ull ^ (uint64_t)y * 0xD8163841ull ^ seed * 0xCB1AB31Full;
    SplitMix s(h); return (double)(s.next() >> 11) * (1.0 / 9007199254740992.0);
}
static double valueNoise(double x, double y, uint64_t seed) {
    double fx = std::floor(x), fy = std::floor(y);
    int64_t ix = (int64_t)fx, iy = (int64_t)fy;
    double tx = x - fx, ty = y - fy;
    tx = tx * tx * (3.0 - 2.0 * tx); ty = ty * ty * (3.0 - 2.0 * ty);
    double a = latticeValue(ix, iy, seed), b = latticeValue(ix + 1, iy, seed);
    double c = latticeValue(ix, iy + 1, seed), d = latticeValue(ix + 1, iy + 1, seed);
    double ab = a + (b - a) * tx, cd = c + (d - c) * tx;
    return ab + (cd - ab) * ty;
}
// fBm in [0, 1): sum of octaves normalised by the total weight
static double fbm(double x, double y, int octaves, uint64_t seed) {
    double s = 0, w = 1, tot = 0, f = 4.0;
    for (int o = 0; o < octaves; o++) { s += w * valueNoise(x * f, y * f, seed + (uint64_t)o * 7919u); tot += w; w *= 0.5; f *= 2.0; }
    return s / tot;
}

void makeTerrain(Mesh& m, int grid, int octaves, float amp, int nBoxes, uint64_t seed, float albedo) {
    m = Mesh();
    m.kd.push_back(v3(albedo, albedo, albedo)); m.matName.push_back("diffuse");
    auto height = [&](double x, double y) { return (float)(amp * (fbm(x, y, octaves, seed) - 0.5)); };
    for (int j = 0; j < grid; j++)
        for (int i = 0; i < grid; i++) {
            double x = (double)i / (grid - 1), y = (double)j / (grid - 1);
            m.v.push_back(v3((float)x, (float)y, height(x, y)));
        }
    for (int j = 0; j + 1 < grid; j++)
        for (int i = 0; i + 1 < grid; i++) {
            int a = j * grid + i, b = a + 1, c = a + grid, d = c + 1;
            m.f.push_back({ a, b, d }); m.fm.push_back(0);
            m.f.push_back({ a, d, c }); m.fm.push_back(0);
        }
    SplitMix r(seed ^ 0xB0C5ull);
    for (int k = 0; k < nBoxes; k++) {
        float cx = r.uniform(0.05f, 0.95f), cy = r.uniform(0.05f, 0.95f);
        Vec3 e = v3(r.uniform(0.01f, 0.05f), r.uniform(0.01f, 0.05f), r.uniform(0.01f, 0.06f));
        float h = height(cx, cy);
        addBox(m, v3(cx - 0.5f * e.x, cy - 0.5f * e.y, h - 0.25f * e.z), v3(cx + 0.5f * e.x, cy + 0.5f * e.y, h + e.z), 0);
    }
}

void terrainCamera(float c2w[16]) { lookAtC2W(v3(0.5f, -0.25f, 0.55f), v3(0.5f, 0.45f, 0.f), v3(0.f, 0.f, 1.f), c2w); }

void makeSphereGrid(Mesh& m, int gridN, int nu, int nv, int gs) {
    m.kd.clear();
    m.kd.push_back(v3(0.5f, 0.5f, 0.5f));
    m.kd.push_back(v3(0.8f, 0.3f, 0.2f));
    m.kd.push_back(v3(0.2f, 0.6f, 0.3f));
    m.kd.push_back(v3(0.25f, 0.35f, 0.8f));
    m.kd.push_back(v3(0.8f, 0.8f, 0.8f));
    const float half = 6.0f;
    int base = (int)m.v.size();
    for (int j = 0; j <= gs; j++)
        for (int i = 0; i <= gs; i++) {
            float x = -half + 2 * half * i / gs, y = -half + 2 * half * j / gs;
            m.v.push_back(v3(x, y, 0.05f * sinf(1.3f * x) * cosf(1.7f * y)));
        }
    for (int j = 0; j < gs; j++)
        for (int i = 0; i < gs; i++) {
            int a = base + j * (gs + 1) + i, b = a + 1, c = a + gs + 1, d = c + 1;
            m.f.push_back({ a, b, d }); m.fm.push_back(0);
            m.f.push_back({ a, d, c }); m.fm.push_back(0);
        }
    float pitch = 2 * half / (gridN + 1), r = pitch * 0.33f;
    for (int sj = 0; sj < gridN; sj++)
        for (int si = 0; si < gridN; si++) {
            int k = sj * gridN + si;
            Vec3 c = v3(-half + pitch * (si + 1), -half + pitch * (sj + 1), r * (1.1f + 0.3f * ((k * 7) % 5) / 4.0f));
            float rr = r * (0.7f + 0.3f * ((k * 3) % 4) / 3.0f);
            int vb = (int)m.v.size();
            for (int j = 0; j <= nv; j++)
                for (int i = 0; i < nu; i++) {
                    float th = 3.14159265f * j / nv, ph = 2 * 3.14159265f * i / nu;
                    float bump = 1.0f + 0.12f * sinf(6 * th + k) * cosf(5 * ph + 0.5f * k);
                    Vec3 d = v3(sinf(th) * cosf(ph), sinf(th) * sinf(ph), cosf(th));
                    m.v.push_back(c + d * (rr * bump));
                }
            int mat = 1 + (k % 4);
            for (int j = 0; j < nv; j++)
                for (int i = 0; i < nu; i++) {
                    int i1 = (i + 1) % nu;
                    int a = vb + j * nu + i, b = vb + j * nu + i1, cc = vb + (j + 1) * nu + i, d = vb + (j + 1) * nu + i1;
                    if (j > 0) { m.f.push_back({ a, cc, b }); m.fm.push_back(mat); }
                    if (j < nv - 1) { m.f.push_back({ b, cc, d }); m.fm.push_back(mat); }
                }
        }
}

void sphereGridCamera(float c2w[16]) { lookAtC2W(v3(0.f, -9.5f, 6.5f), v3(0.f, 0.f, 0.4f), v3(0, 0, 1), c2w); }

void makeIcosphere(Mesh& m, int level, int mat) {
    const float t = 1.6180339887f;
    std::vector<Vec3> v = { v3(-1, t, 0), v3(1, t, 0), v3(-1, -t, 0), v3(1, -t, 0), v3(0, -1, t), v3(0, 1, t),
                            v3(0, -1, -t), v3(0, 1, -t), v3(t, 0, -1), v3(t, 0, 1), v3(-t, 0, -1), v3(-t, 0, 1) };
    std::vector<Tri> f = { { 0, 11, 5 }, { 0, 5, 1 }, { 0, 1, 7 }, { 0, 7, 10 }, { 0, 10, 11 }, { 1, 5, 9 }, { 5, 11, 4 },
                           { 11, 10, 2 }, { 10, 7, 6 }, { 7, 1, 8 }, { 3, 9, 4 }, { 3, 4, 2 }, { 3, 2, 6 }, { 3, 6, 8 },
                           { 3, 8, 9 }, { 4, 9, 5 }, { 2, 4, 11 }, { 6, 2, 10 }, { 8, 6, 7 }, { 9, 8, 1 } };
    for (auto& p : v) p = normalize(p);
    for (int l = 0; l < level; l++) {
        std::map<std::pair<int, int>, int> mid;
        auto midpoint = [&](int a, int b) {
            auto key = std::make_pair(a < b ? a : b, a < b ? b : a);
            auto it = mid.find(key); if (it != mid.end()) return it->second;
            v.push_back(normalize((v[a] + v[b]) * 0.5f)); int id = (int)v.size() - 1; mid[key] = id; return id;
        };
        std::vector<Tri> g;
        for (const Tri& q : f) {
            int ab = midpoint(q.a, q.b), bc = midpoint(q.b, q.c), ca = midpoint(q.c, q.a);
            g.push_back({ q.a, ab, ca }); g.push_back({ q.b, bc, ab }); g.push_back({ q.c, ca, bc }); g.push_back({ ab, bc, ca });
        }
        f.swap(g);
    }
    int base = (int)m.v.size();
    for (const Vec3& p : v) m.v.push_back(p);
    for (const Tri& q : f) { m.f.push_back({ q.a + base, q.b + base, q.c + base }); m.fm.push_back(mat); }
}

void assignGeneratedMaterials(Mesh& m, int first, int n, uint64_t seed) {
    if (m.kd.empty()) m.kd.push_back(v3(0.8f, 0.8f, 0.8f));
    std::vector<Material> mats(1 + (size_t)n);
    Material g = defaultMaterial();   // glossy-diffuse ground / terrain
    g.baseColor = m.kd[0]; g.roughness = 0.5f;
    mats[0] = g;
    const Vec3 metals[4][2] = { { v3(1.00f, 0.78f, 0.34f), v3(0.95f, 0.86f, 0.55f) },    // gold, F82 tint
                                { v3(0.95f, 0.64f, 0.54f), v3(0.93f, 0.75f, 0.70f) },    // copper
                                { v3(0.91f, 0.92f, 0.92f), v3(0.96f, 0.97f, 0.98f) },    // aluminium
                                { v3(0.97f, 0.96f, 0.91f), v3(1.00f, 1.00f, 1.00f) } };  // silver
    SplitMix r(seed * 0x9E3779B97F4A7C15ull + 0x632BE59BD9B4E019ull);
    for (int b = 0; b < n; b++) {
        Material x = defaultMaterial();
        const float kind = r.next01(), rough = r.uniform(0.05f, 0.6f);
        if (kind < 0.30f) {
            const int k = (int)(r.next01() * 4.f) & 3;
            x.metallic = 1.f; x.baseColor = metals[k][0]; x.specularColor = metals[k][1]; x.roughness = rough;
        } else if (kind < 0.55f) {
            x.transmission = 1.f; x.ior = 1.5f; x.roughness = rough;
            x.baseColor = v3(r.uniform(0.8f, 1.f), r.uniform(0.8f, 1.f), r.uniform(0.8f, 1.f));
        } else if (kind < 0.85f) {
            x.baseColor = v3(r.uniform(0.05f, 0.9f), r.uniform(0.05f, 0.9f), r.uniform(0.05f, 0.9f)); x.roughness = rough;
        } else {
            x = lambertMaterial(v3(0.5f, 0.5f, 0.5f));
            x.emission = v3(r.uniform(0.6f, 1.f), r.uniform(0.6f, 1.f), r.uniform(0.6f, 1.f)) * 8.f;
        }
        mats[1 + b] = x;
    }
    if (m.fm.empty()) m.fm.assign(m.f.size(), 0);
    for (int b = 0; b < n; b++)
        for (int t = 0; t < 12; t++) {
            const size_t tri = (size_t)first + 12 * (size_t)b + (size_t)t;
            if (tri < m.fm.size()) m.fm[tri] = 1 + b;
        }
    m.kd.resize(mats.size(), v3(0.8f, 0.8f, 0.8f));
    for (size_t j = 0; j < mats.size(); j++) m.kd[j] = mats[j].baseColor;
    m.mats = mats;
}

void addGeneratedTextures(Mesh& m, int size, uint64_t seed) {
    const float repeat = 8.f;
    m.uv.assign(6 * m.f.size(), 0.f);
    for (size_t t = 0; t < m.f.size(); t++) {
        const Vec3 p[3] = { m.v[m.f[t].a], m.v[m.f[t].b], m.v[m.f[t].c] };
        const Vec3 n = cross(p[1] - p[0], p[2] - p[0]);
        const float ax = std::fabs(n.x), ay = std::fabs(n.y), az = std::fabs(n.z);
        for (int c = 0; c < 3; c++) {   // project out the dominant axis
            const float u = az >= ax && az >= ay ? p[c].x : p[c].y, v = az >= ax && az >= ay ? p[c].y : p[c].z;
            m.uv[6 * t + 2 * c] = u * repeat; m.uv[6 * t + 2 * c + 1] = v * repeat;
        }
    }
    SplitMix rng(seed * 0x9E3779B97F4A7C15ull + 17);
    const int cells = 8;
    std::vector<float> noise((size_t)(cells + 1) * (cells + 1));
    for (float& x : noise) x = rng.next01();
    auto smooth = [&](float x, float y) {   // bilinear value noise over the cell lattice, periodic
        const float gx = x * cells, gy = y * cells;
        const int ix = (int)gx % cells, iy = (int)gy % cells;
        const float fx = gx - std::floor(gx), fy = gy - std::floor(gy);
        auto at = [&](int i, int j) { return noise[(size_t)(j % cells) * (cells + 1) + (i % cells)]; };
        return (at(ix, iy) * (1 - fx) + at(ix + 1, iy) * fx) * (1 - fy) + (at(ix, iy + 1) * (1 - fx) + at(ix + 1, iy + 1) * fx) * fy;
    };
    std::vector<int> metalCell((size_t)cells * cells);
    for (int& c : metalCell) c = rng.next01() < 0.25f ? 1 : 0;
    auto byte = [](float x) { return (uint32_t)std::lround(std::fmin(std::fmax(x, 0.f), 1.f) * 255.f); };
    TextureImage base, normal, packed;
    for (TextureImage* T : { &base, &normal, &packed }) { T->w = T->h = size; T->rgba.resize((size_t)size * size); }
    base.path = "<generated base colour>"; base.srgb = true;
    normal.path = "<generated normal map>"; packed.path = "<generated roughness / metalness>";
    const float k = 6.2831853f * 4.f, amp = 0.35f;
    for (int y = 0; y < size; y++)
        for (int x = 0; x < size; x++) {
            const float u = (x + 0.5f) / size, v = (y + 0.5f) / size;
            const int cx = (int)(u * cells), cy = (int)(v * cells);
            const float nz = smooth(u, v);
            const bool odd = ((cx + cy) & 1) != 0;
            const Vec3 c0 = v3(0.85f, 0.72f, 0.55f), c1 = v3(0.25f, 0.35f, 0.45f);
            const Vec3 col = (odd ? c1 : c0) * (0.7f + 0.3f * nz);
            base.rgba[(size_t)y * size + x] = byte(col.x) | (byte(col.y) << 8) | (byte(col.z) << 16) | (255u << 24);
            // bump h = amp / k (sin(k u) + sin(k v)): slopes amp cos(k u), amp cos(k v)
            const Vec3 nn = normalize(v3(-amp * std::cos(k * u), -amp * std::cos(k * v), 1.f));
            normal.rgba[(size_t)y * size + x] = byte(nn.x * 0.5f + 0.5f) | (byte(nn.y * 0.5f + 0.5f) << 8) |
                                                (byte(nn.z * 0.5f + 0.5f) << 16) | (255u << 24);
            const float rough = 0.3f + 0.7f * nz;
            packed.rgba[(size_t)y * size + x] = (byte(rough) << 8) | ((metalCell[(size_t)cy * cells + cx] ? 255u : 0u) << 16) | (255u << 24);
        }
    m.tex = { base, normal, packed };
    const size_t nm = std::max(m.kd.size(), std::max(m.mats.size(), (size_t)1));
    MaterialMaps x = noMaps();
    x.baseColor = 0; x.normal = 1; x.normalScale = 1.f; x.roughness = 2; x.roughnessChannel = 1; x.metallic = 2; x.metallicChannel = 2;
    m.maps.assign(nm, x);
}

}  // namespace cwp
