// core/vecmath.h - 3- and 4-component float vectors shared by the C++17 host code and the gfx950 kernels.
//
// Plain aggregates (no HIP vector types), so the same header compiles with g++ for the host library and with hipcc
// for device code. Every operator spells out its IEEE operation order: the build compiles with -ffp-contract=off, and
// places that want a fused multiply-add call fmaf() explicitly, so the host and device evaluate identically.
// Arithmetic follows the reference's src/core/vec.h:6-25 (component-wise ops, dot as ((x*x + y*y) + z*z), normalize by
// the reciprocal of the length).
#pragma once
#include <math.h>
#include <stdint.h>
#include <string.h>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define CWP_HD __host__ __device__ __forceinline__
#else
#define CWP_HD inline
#endif

namespace cwp {

struct Vec3 { float x, y, z; };
struct alignas(16) Vec4 { float x, y, z, w; };

CWP_HD Vec3 v3(float x, float y, float z) { Vec3 r; r.x = x; r.y = y; r.z = z; return r; }
CWP_HD Vec4 v4(float x, float y, float z, float w) { Vec4 r; r.x = x; r.y = y; r.z = z; r.w = w; return r; }
CWP_HD Vec4 v4(Vec3 a, float w) { return v4(a.x, a.y, a.z, w); }
CWP_HD Vec3 xyz(Vec4 a) { return v3(a.x, a.y, a.z); }

CWP_HD Vec3 operator+(Vec3 a, Vec3 b) { return v3(a.x + b.x, a.y + b.y, a.z + b.z); }
CWP_HD Vec3 operator-(Vec3 a, Vec3 b) { return v3(a.x - b.x, a.y - b.y, a.z - b.z); }
CWP_HD Vec3 operator-(Vec3 a) { return v3(-a.x, -a.y, -a.z); }
CWP_HD Vec3 operator*(Vec3 a, Vec3 b) { return v3(a.x * b.x, a.y * b.y, a.z * b.z); }
CWP_HD Vec3 operator*(Vec3 a, float s) { return v3(a.x * s, a.y * s, a.z * s); }
CWP_HD Vec3 operator*(float s, Vec3 a) { return v3(a.x * s, a.y * s, a.z * s); }
CWP_HD float dot(Vec3 a, Vec3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
CWP_HD Vec3 cross(Vec3 a, Vec3 b) { return v3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x); }
CWP_HD float length(Vec3 a) { return sqrtf(dot(a, a)); }
CWP_HD Vec3 normalize(Vec3 a) { float l = length(a); return l > 0.f ? a * (1.0f / l) : a; }
CWP_HD Vec3 vmin(Vec3 a, Vec3 b) { return v3(fminf(a.x, b.x), fminf(a.y, b.y), fminf(a.z, b.z)); }
CWP_HD Vec3 vmax(Vec3 a, Vec3 b) { return v3(fmaxf(a.x, b.x), fmaxf(a.y, b.y), fmaxf(a.z, b.z)); }
CWP_HD float maxc(Vec3 a) { return fmaxf(a.x, fmaxf(a.y, a.z)); }
CWP_HD float comp(Vec3 a, int k) { return k == 0 ? a.x : (k == 1 ? a.y : a.z); }
CWP_HD float comp(const Vec4& a, int k) { return k == 0 ? a.x : (k == 1 ? a.y : (k == 2 ? a.z : a.w)); }

// bit casts
CWP_HD float asFloat(int i) { float f; memcpy(&f, &i, 4); return f; }
CWP_HD float asFloatU(uint32_t u) { float f; memcpy(&f, &u, 4); return f; }
CWP_HD int asInt(float f) { int i; memcpy(&i, &f, 4); return i; }
CWP_HD uint32_t asUint(float f) { uint32_t u; memcpy(&u, &f, 4); return u; }

}  // namespace cwp
