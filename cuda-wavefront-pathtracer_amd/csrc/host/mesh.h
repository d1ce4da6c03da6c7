// host/mesh.h - indexed triangle meshes, boxes, and the procedural scenes used by the CLI, the tests and the bench.
#pragma once
#include "core/material.h"
#include "core/texture.h"
#include "core/vecmath.h"
#include <cfloat>
#include <string>
#include <vector>

namespace cwp {

struct Tri { int a, b, c; };

// An RGBA8 texture image (rows top first, r in the low byte of each texel).
struct TextureImage {
    std::string path;
    int w = 0, h = 0;
    bool srgb = false;             // colour data (base colour, emission): decoded from sRGB on lookup
    std::vector<uint32_t> rgba;
};

// Indexed triangle mesh with a per-triangle material id and a per-material Lambert albedo table
// (reference: src/scene/mesh.h:9-26; material names come from OBJ usemtl).
struct Mesh {
    std::vector<Vec3> v;       // vertex positions
    std::vector<Vec3> n;       // optional per-vertex normals (same size as v) or empty
    std::vector<Tri> f;        // triangles
    std::vector<int> fm;       // per-triangle material id (empty: all 0)
    std::vector<Vec3> kd;      // albedo per material id
    std::vector<std::string> matName;
    std::vector<Material> mats; // full materials per id (from a materials file or a generator); empty: Lambert kd
    std::vector<float> uv;      // texture coordinates: 6 floats per triangle (u, v of its 3 corners) or empty
    std::vector<MaterialMaps> maps;   // texture bindings per material id (empty: none)
    std::vector<TextureImage> tex;    // the images the bindings name
    bool hasN() const { return !n.empty() && n.size() == v.size(); }
    bool hasUV() const { return !f.empty() && uv.size() == 6 * f.size(); }
    int mat(size_t t) const { return fm.empty() ? 0 : fm[t]; }
};

// Axis-aligned box; empty boxes are inverted (reference: src/accel/aabb.h:6-13).
struct AABB {
    Vec3 lo = v3(FLT_MAX, FLT_MAX, FLT_MAX), hi = v3(-FLT_MAX, -FLT_MAX, -FLT_MAX);
    void grow(Vec3 p) { lo = vmin(lo, p); hi = vmax(hi, p); }
    void grow(const AABB& b) { lo = vmin(lo, b.lo); hi = vmax(hi, b.hi); }
    // surface area 2(dx dy + dy dz + dz dx); 0 for an empty box
    float area() const {
        if (hi.x < lo.x) return 0.f;
        Vec3 d = hi - lo;
        return 2.f * (d.x * d.y + d.y * d.z + d.z * d.x);
    }
    Vec3 extent() const { return hi - lo; }
    Vec3 center() const { return (lo + hi) * 0.5f; }
};

AABB triBox(const Mesh& m, size_t t);
AABB meshBox(const Mesh& m);

// Appends an axis-aligned box (12 triangles, outward winding) with material mat.
void addBox(Mesh& m, Vec3 lo, Vec3 hi, int mat);

// Seeded splitmix64 stream; next01() gives a float in [0, 1) with 24 random bits.
struct SplitMix {
    uint64_t s;
    explicit SplitMix(uint64_t seed) : s(seed) {}
    uint64_t next() {
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull; z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    float next01() { return (float)(next() >> 40) * (1.0f / 16777216.0f); }
    float uniform(float a, float b) { return a + (b - a) * next01(); }
};

// Box-cloud scene: nBoxes axis-aligned boxes, centres uniform in [0,1]^3, edge lengths uniform in [0.01, 0.08], plus a
// 2-triangle ground square [-1,2]^2 at z = 0; one material of albedo `albedo`.
void makeBoxCloud(Mesh& m, int nBoxes, uint64_t seed, float albedo);
// Default view of the box cloud (row-major camera-to-world, OpenGL camera axes, z-up world).
void boxCloudCamera(float c2w[16]);

// Terrain scene: grid x grid vertex fBm heightfield over [0,1]^2 (value noise, `octaves` octaves, lacunarity 2, gain 0.5,
// peak-to-peak amplitude `amp`), 2 (grid-1)^2 triangles, plus nBoxes boxes resting on it; one material of albedo `albedo`.
void makeTerrain(Mesh& m, int grid, int octaves, float amp, int nBoxes, uint64_t seed, float albedo);
void terrainCamera(float c2w[16]);

// The reference's built-in `--synthetic N` scene: ground (gs x gs quads, gently waved) and N x N bumpy spheres of
// nu x nv quads (reference: src/scene/synthetic.cpp:6-50; its default view: src/scene/synthetic.cpp:50).
void makeSphereGrid(Mesh& m, int gridN, int nu, int nv, int gs);
void sphereGridCamera(float c2w[16]);

// OpenPBR-subset materials for a generated scene's boxes (triangles [first, first + 12 n)): material 0 (terrain or
// ground) becomes glossy-diffuse; each box gets its own material drawn from the seed: rough conductor (F0 of gold,
// copper, aluminium or silver, F82 tint, roughness U[0.05, 0.6]) 30 %, rough dielectric (IOR 1.5, roughness
// U[0.05, 0.6]) 25 %, glossy-diffuse (random colour, roughness U[0.05, 0.6]) 30 %, emitter (Lambert 0.5, emission
// strength 8) 15 %.
void assignGeneratedMaterials(Mesh& m, int first, int n, uint64_t seed);

// Generated textures for a procedural scene: texture coordinates by planar projection along each triangle's dominant
// normal axis (8 repeats per unit), and three size x size images bound to every material: a base-colour checker
// (8 x 8 cells, value-noise modulated, sRGB), a tangent-space normal map of a sinusoidal bump field, and a packed map
// with roughness U[0.3, 1] in G and metalness 0 / 1 per cell in B (glTF channels).
void addGeneratedTextures(Mesh& m, int size, uint64_t seed);

// Subdivided icosahedron of radius 1 at the origin (20 * 4^level triangles).
void makeIcosphere(Mesh& m, int level, int mat);

}  // namespace cwp
