// host/instances.h - two-level scenes: meshes, instances with per-frame transforms, and the per-frame instance records
// the renderer's TLAS is built from.
//
// Scene file (JSON; reference format, src/scene/instances.h:1-16):
// {
//   "frames": 30,
//   "meshes": [ {"file": "a.obj", "yup_to_zup": 1}, ... ],
//   "instances": [
//     {"mesh": 0, "transform": [12 or 16 numbers, row-major object-to-world]},
//     {"mesh": 1, "translate": [x,y,z], "axis": [x,y,z], "angle_deg": a, "scale": s or [sx,sy,sz],
//                 "translate_per_frame": [..], "angle_deg_per_frame": da, "scale_per_frame": [..]},
//     {"mesh": 1, "transforms": [[12 or 16], [..], ...]}       // frame f uses min(f, n - 1)
//   ]
// }
// Parametric instances: M(f) = T(translate + f v_t) R(axis, angle + f v_a) S(scale + f v_s), evaluated in double.
// Mesh paths are relative to the scene file. A mesh's "materials" key (full-integrator material files) is ignored:
// this renderer shades with the meshes' Lambert albedos.
#pragma once
#include "core/vecmath.h"
#include "device/traverse_types.h"
#include "host/mesh.h"
#include <cstdint>
#include <string>
#include <vector>

namespace cwp {

struct Xform { float m[12]; };   // row-major 3x4 object-to-world

struct InstanceDesc {
    int mesh = 0;
    bool hasMatrix = false;
    Xform matrix{};
    std::vector<Xform> perFrame;
    Vec3 translate = v3(0, 0, 0), translateVel = v3(0, 0, 0), axis = v3(0, 0, 1), scale = v3(1, 1, 1), scaleVel = v3(0, 0, 0);
    float angleDeg = 0.f, angleVel = 0.f;
    // generated scenes: a rigid transform re-drawn every frame from (randomKey, frame)
    bool random = false;
    uint64_t randomKey = 0;
    Vec3 spanLo = v3(0, 0, 0), spanHi = v3(0, 0, 0);   // translation range of random instances
};
struct MeshDesc { std::string file; int yup = -1; std::string materials; };   // materials: per-mesh materials file
struct SceneDesc { std::vector<MeshDesc> meshes; std::vector<InstanceDesc> instances; int frames = 1; };

bool loadSceneFile(const std::string& path, SceneDesc& S, std::string& err);
// Object-to-world transform of an instance at a frame (reference: instanceXform, src/scene/instances.cpp:48-63).
Xform instanceXform(const InstanceDesc& I, int frame);
// Inverse in double precision and the sign of the determinant of the linear part (reference: invertXform,
// src/scene/instances.cpp:65-79).
void invertXform(const Xform& M, Xform& inv, float& detSign);
inline Vec3 xfPoint(const Xform& X, Vec3 p) {
    const float* m = X.m;
    return v3(m[0] * p.x + m[1] * p.y + m[2] * p.z + m[3], m[4] * p.x + m[5] * p.y + m[6] * p.z + m[7],
              m[8] * p.x + m[9] * p.y + m[10] * p.z + m[11]);
}
inline Vec3 xfVector(const Xform& X, Vec3 p) {
    const float* m = X.m;
    return v3(m[0] * p.x + m[1] * p.y + m[2] * p.z, m[4] * p.x + m[5] * p.y + m[6] * p.z, m[8] * p.x + m[9] * p.y + m[10] * p.z);
}
// normal transform with the inverse: n' = inv^T n (not normalised)
inline Vec3 xfNormal(const Xform& inv, Vec3 n) {
    const float* m = inv.m;
    return v3(m[0] * n.x + m[4] * n.y + m[8] * n.z, m[1] * n.x + m[5] * n.y + m[9] * n.z, m[2] * n.x + m[6] * n.y + m[10] * n.z);
}

// The two-level scene of the reference's H-PLOC configuration: 5 procedural meshes (icospheres and fBm heightfields of
// about 1k, 10k, 50k, 100k and 200k triangles) plus a ground square, and nInst instances with rigid transforms redrawn
// every frame (uniform random rotation, translation uniform in a 24 x 24 x 2 box), seeded. small: the same scene with
// meshes of about 300, 700, 3k, 1.3k and 7k triangles (quick tests).
void makeInstanceScene(std::vector<Mesh>& meshes, SceneDesc& S, int nInst, uint64_t seed, bool small = false);
void instanceSceneCamera(float c2w[16]);

// Appends the geometry of `src` transformed by X (positions by X, normals by inv^T, unnormalised) with material ids
// offset by matOffset: the flattened equivalent of one instance (reference: appendGeometry, instances.cpp:100-122).
void appendTransformed(Mesh& dst, const Mesh& src, int matOffset, const Xform& X);

}  // namespace cwp
