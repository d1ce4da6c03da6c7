// device/traverse.h - gfx950 ray traversal: BVH2 (64 B nodes) and compressed BVH8 (80 B nodes), closest and any hit,
// with three triangle tests. One ray per lane; the traversal stack lives in LDS (SL entries per lane, lane-minor so a
// wave touching the same depth hits consecutive banks) and spills to a per-thread global slab beyond SL entries.
//
// Algorithms (restated from the reference, src/accel/traverse.cuh):
//  * BVH2 (lines 45-70): both child slabs per node, leaves intersected on the spot, nearer child first, farther pushed.
//  * BVH8 (lines 112-152): per node decode the per-axis power-of-two scales, test the 8 quantised child slabs, intersect
//    hit leaves in slot order, record internal hits in octant-sorted bit order (slot ^ ray octant), then continue with
//    the lowest set bit; the remaining group (child base, imask, hit bits) goes on the stack.
//  * triangles: Moller-Trumbore (lines 13-23), Woop unit-triangle rows (lines 24-38, t first), and the watertight test
//    of Woop, Benthin and Wald (JCGT 2013) with its double-precision fallback on exact edge hits.
// Every expression spells out its fused multiply-adds (the build uses -ffp-contract=off), so the host oracle evaluates
// the same floating-point operations and finds the same hits.
#pragma once
#include "core/vecmath.h"
#include "device/traverse_types.h"

namespace cwp {

struct Hit { float t, u, v; int tri; int inst; };   // inst: instance of the hit in two-level scenes

// Device stride of a compressed BVH8 node in float4: the 80-byte node padded to one 128-byte cache line (deviceBVH8 on
// the host writes it; the child base carries the leaf-slot mask in its top byte).
constexpr int kBvh8Stride = 8;

// Per-ray constants shared by every node and triangle test.
struct RayPre {
    Vec3 o, d;
    float tmin;
    float ix, iy, iz;          // reciprocal direction, |d| clamped to 1e-20
    uint32_t oct;              // bit k set: direction component k > 0 (child visiting order)
    uint32_t pos;              // bit k set: reciprocal direction component k >= 0 (near slab plane is the low one)
    int kx, ky, kz;            // watertight: dominant axis kz and the two others
    float Sx, Sy, Sz;          // watertight shear
};

__device__ __forceinline__ float safeRcp(float a) { const float e = 1e-20f; return 1.f / (fabsf(a) > e ? a : copysignf(e, a)); }

template <int TRI>
__device__ __forceinline__ RayPre prepareRay(Vec3 o, Vec3 d, float tmin) {
    RayPre r;
    r.o = o; r.d = d; r.tmin = tmin;
    r.ix = safeRcp(d.x); r.iy = safeRcp(d.y); r.iz = safeRcp(d.z);
    r.oct = (d.x > 0.f ? 1u : 0u) | (d.y > 0.f ? 2u : 0u) | (d.z > 0.f ? 4u : 0u);
    r.pos = (r.ix >= 0.f ? 1u : 0u) | (r.iy >= 0.f ? 2u : 0u) | (r.iz >= 0.f ? 4u : 0u);
    r.kx = r.ky = r.kz = 0; r.Sx = r.Sy = r.Sz = 0.f;
    if (TRI == kTriWatertight) {
        float ax = fabsf(d.x), ay = fabsf(d.y), az = fabsf(d.z);
        int kz = (ax > ay) ? (ax > az ? 0 : 2) : (ay > az ? 1 : 2);
        int kx = kz == 2 ? 0 : kz + 1, ky = kx == 2 ? 0 : kx + 1;
        if (comp(d, kz) < 0.f) { int t = kx; kx = ky; ky = t; }
        r.kx = kx; r.ky = ky; r.kz = kz;
        float dz = comp(d, kz);
        r.Sx = comp(d, kx) / dz; r.Sy = comp(d, ky) / dz; r.Sz = 1.f / dz;
    }
    return r;
}

__device__ __forceinline__ Vec4 ldg4(const Vec4* p) { return *p; }

// ------------------------------------------------------------------------------------------------ triangle tests
// Tests triangle position t (leaf order). Updates the hit and returns true if it is closer than hit.t.
template <int TRI>
__device__ __forceinline__ bool triTest(const Vec4* __restrict__ tri, int t, const RayPre& r, Hit& h) {
    if (TRI == kTriWoop) {
        // the three rows are read together (one memory latency per triangle instead of three in sequence)
        const Vec4 r2 = ldg4(tri + 3 * t + 2), r0 = ldg4(tri + 3 * t), r1 = ldg4(tri + 3 * t + 1);
        float Oz = fmaf(r.o.z, r2.z, fmaf(r.o.y, r2.y, fmaf(r.o.x, r2.x, r2.w)));
        float Dz = fmaf(r.d.z, r2.z, fmaf(r.d.y, r2.y, r.d.x * r2.x));
        float tt = -Oz / Dz;
        if (!(tt > r.tmin && tt < h.t)) return false;
        float Ox = fmaf(r.o.z, r0.z, fmaf(r.o.y, r0.y, fmaf(r.o.x, r0.x, r0.w)));
        float Dx = fmaf(r.d.z, r0.z, fmaf(r.d.y, r0.y, r.d.x * r0.x));
        float u = fmaf(tt, Dx, Ox);
        if (!(u >= 0.f && u <= 1.f)) return false;
        float Oy = fmaf(r.o.z, r1.z, fmaf(r.o.y, r1.y, fmaf(r.o.x, r1.x, r1.w)));
        float Dy = fmaf(r.d.z, r1.z, fmaf(r.d.y, r1.y, r.d.x * r1.x));
        float v = fmaf(tt, Dy, Oy);
        if (!(v >= 0.f && u + v <= 1.f)) return false;
        h.t = tt; h.u = u; h.v = v; h.tri = t;
        return true;
    } else if (TRI == kTriMT) {
        Vec4 a = ldg4(tri + 3 * t), b = ldg4(tri + 3 * t + 1), c = ldg4(tri + 3 * t + 2);
        Vec3 v0 = xyz(a), e1 = xyz(b) - v0, e2 = xyz(c) - v0;
        Vec3 pv = cross(r.d, e2);
        float det = dot(e1, pv);
        if (fabsf(det) < 1e-12f) return false;
        float inv = 1.f / det;
        Vec3 tv = r.o - v0;
        float u = dot(tv, pv) * inv;
        if (u < 0.f || u > 1.f) return false;
        Vec3 qv = cross(tv, e1);
        float v = dot(r.d, qv) * inv;
        if (v < 0.f || u + v > 1.f) return false;
        float tt = dot(e2, qv) * inv;
        if (!(tt > r.tmin && tt < h.t)) return false;
        h.t = tt; h.u = u; h.v = v; h.tri = t;
        return true;
    } else {
        Vec4 pa = ldg4(tri + 3 * t), pb = ldg4(tri + 3 * t + 1), pc = ldg4(tri + 3 * t + 2);
        Vec3 A = xyz(pa) - r.o, B = xyz(pb) - r.o, C = xyz(pc) - r.o;
        float Akz = comp(A, r.kz), Bkz = comp(B, r.kz), Ckz = comp(C, r.kz);
        float Ax = comp(A, r.kx) - r.Sx * Akz, Ay = comp(A, r.ky) - r.Sy * Akz;
        float Bx = comp(B, r.kx) - r.Sx * Bkz, By = comp(B, r.ky) - r.Sy * Bkz;
        float Cx = comp(C, r.kx) - r.Sx * Ckz, Cy = comp(C, r.ky) - r.Sy * Ckz;
        float U = Cx * By - Cy * Bx, V = Ax * Cy - Ay * Cx, W = Bx * Ay - By * Ax;
        if (U == 0.f || V == 0.f || W == 0.f) {   // exact edge or vertex hit: re-evaluate the edge functions in double
            U = (float)((double)Cx * (double)By - (double)Cy * (double)Bx);
            V = (float)((double)Ax * (double)Cy - (double)Ay * (double)Cx);
            W = (float)((double)Bx * (double)Ay - (double)By * (double)Ax);
        }
        if ((U < 0.f || V < 0.f || W < 0.f) && (U > 0.f || V > 0.f || W > 0.f)) return false;
        float det = U + V + W;
        if (det == 0.f) return false;
        float Az = r.Sz * Akz, Bz = r.Sz * Bkz, Cz = r.Sz * Ckz;
        float T = U * Az + V * Bz + W * Cz;
        float rcpDet = 1.f / det;
        float tt = T * rcpDet;
        if (!(tt > r.tmin && tt < h.t)) return false;
        h.t = tt; h.u = V * rcpDet; h.v = W * rcpDet; h.tri = t;
        return true;
    }
}

// ------------------------------------------------------------------------------------------------ traversal stack
// LDS pointers are typed as address space 3, so every stack access is a ds_read/ds_write (never a flat access that
// also waits on vector memory), whatever the optimiser does with the loop around it.
template <class T> using LdsPtr = __attribute__((address_space(3))) T*;

// Per-lane traversal stack: the first SL entries in LDS, lane-minor (entry j of lane l at lds[j * kStackStride + l],
// conflict-free), deeper entries in the lane's HBM spill area.
constexpr int kStackStride = 64;   // lanes per workgroup: one wave
template <typename E, int SL>
struct LaneStack {
    LdsPtr<E> lds; E* spill; int sp;
    int floor = 0;   // entries below this belong to an enclosing traversal (the TLAS, while a BLAS is traversed)
    __device__ __forceinline__ void push(E e) {
        if (SL > 0 && sp < SL) lds[sp * kStackStride] = e; else spill[sp - SL] = e;
        sp++;
    }
    __device__ __forceinline__ E pop() {
        sp--;
        if (SL > 0 && sp < SL) return lds[sp * kStackStride];
        return spill[sp - SL];
    }
    // entry e (0 = bottom) without popping it
    __device__ __forceinline__ E at(int e) const { return (SL > 0 && e < SL) ? lds[e * kStackStride] : spill[e - SL]; }
};

// BVH8 stack entry: child base in the low word, (internal-child mask << 8) | remaining hit bits in the high word
using Group = uint64_t;
__device__ __forceinline__ Group makeGroup(uint32_t base, uint32_t bits) { return ((uint64_t)bits << 32) | base; }

// ------------------------------------------------------------------------------------------------ traversal state
// Resumable state of one ray's traversal: the node to visit next and, for BVH8, the group of hit children still to
// visit (child base, internal-child mask, octant-ordered hit bits). With the lane's stack this is everything a
// persistent kernel needs to suspend a ray between node visits.
struct TravState { int node; uint32_t gBase, gMask, gHits; };

__device__ __forceinline__ void travBegin(TravState& s, int root) { s.node = root; s.gBase = 0; s.gMask = 0; s.gHits = 0; }

// ------------------------------------------------------------------------------------------------ BVH2
// One node visit. Returns true when the traversal is over (stack exhausted, or for ANY a hit was found).
template <bool ANY, int TRI, class Stack>
__device__ __forceinline__ bool stepBVH2(const Vec4* __restrict__ nodes, const Vec4* __restrict__ tri, const RayPre& r,
                                         Hit& h, Stack& st, TravState& ts) {
    const float odx = r.o.x * r.ix, ody = r.o.y * r.iy, odz = r.o.z * r.iz;
    const Vec4* nb = nodes + (size_t)ts.node * 4;
    const Vec4 n0 = ldg4(nb), n1 = ldg4(nb + 1), nz = ldg4(nb + 2), mt = ldg4(nb + 3);
    int c0 = asInt(mt.x), c1 = asInt(mt.y);
    const int k0 = asInt(mt.z), k1 = asInt(mt.w);
    float a0 = fmaf(n0.x, r.ix, -odx), a1 = fmaf(n0.y, r.ix, -odx), a2 = fmaf(n0.z, r.iy, -ody), a3 = fmaf(n0.w, r.iy, -ody);
    float a4 = fmaf(nz.x, r.iz, -odz), a5 = fmaf(nz.y, r.iz, -odz);
    float b0 = fmaf(n1.x, r.ix, -odx), b1 = fmaf(n1.y, r.ix, -odx), b2 = fmaf(n1.z, r.iy, -ody), b3 = fmaf(n1.w, r.iy, -ody);
    float b4 = fmaf(nz.z, r.iz, -odz), b5 = fmaf(nz.w, r.iz, -odz);
    float t0n = fmaxf(fmaxf(fminf(a0, a1), fminf(a2, a3)), fmaxf(fminf(a4, a5), r.tmin));
    float t0f = fminf(fminf(fmaxf(a0, a1), fmaxf(a2, a3)), fminf(fmaxf(a4, a5), h.t));
    float t1n = fmaxf(fmaxf(fminf(b0, b1), fminf(b2, b3)), fmaxf(fminf(b4, b5), r.tmin));
    float t1f = fminf(fminf(fmaxf(b0, b1), fmaxf(b2, b3)), fminf(fmaxf(b4, b5), h.t));
    bool h0 = t0n <= t0f, h1 = t1n <= t1f;
    if (h0 && k0 > 0) {
        for (int t = c0; t < c0 + k0; t++) if (triTest<TRI>(tri, t, r, h) && ANY) return true;
        h0 = false;
    }
    if (h1 && k1 > 0) {
        for (int t = c1; t < c1 + k1; t++) if (triTest<TRI>(tri, t, r, h) && ANY) return true;
        h1 = false;
    }
    if (!h0 && !h1) {
        if (st.sp == st.floor) return true;
        ts.node = st.pop();
    } else if (h0 && h1) {
        if (t1n < t0n) { int x = c0; c0 = c1; c1 = x; }
        st.push(c1);
        ts.node = c0;
    } else ts.node = h0 ? c0 : c1;
    return false;
}

template <bool ANY, int TRI, class Stack>
__device__ __forceinline__ bool traceBVH2(const Vec4* __restrict__ nodes, const Vec4* __restrict__ tri, int root,
                                          const RayPre& r, Hit& h, Stack& st) {
    TravState ts; travBegin(ts, root);
    while (!stepBVH2<ANY, TRI>(nodes, tri, r, h, st, ts)) {}
    return h.tri >= 0;
}

// ------------------------------------------------------------------------------------------------ BVH4
// 8 float4 per node (128 B, one cache line): lo.x, hi.x, lo.y, hi.y, lo.z, hi.z of the 4 children (child j in
// component j), child int4 (node index; first triangle when the count is > 0; -1 = empty slot), count int4.
// One node visit: the 4 slabs against the hit distance at entry (near / far planes chosen by the ray's direction
// signs: bit-identical to min / max), the hit leaves in slot order, then the internal hits that can still hold a closer
// hit, nearest first (sorted by entry distance with a 5-comparator network; the far ones are pushed farthest first).
template <bool ANY, int TRI, class Stack>
__device__ __forceinline__ bool stepBVH4(const Vec4* __restrict__ nodes, const Vec4* __restrict__ tri, const RayPre& r,
                                         Hit& h, Stack& st, TravState& ts) {
    const float odx = r.o.x * r.ix, ody = r.o.y * r.iy, odz = r.o.z * r.iz;
    const Vec4* nb = nodes + (size_t)ts.node * 8;
    const Vec4 lx = ldg4(nb), hx = ldg4(nb + 1), ly = ldg4(nb + 2), hy = ldg4(nb + 3), lz = ldg4(nb + 4), hz = ldg4(nb + 5);
    const Vec4 cf = ldg4(nb + 6), kf = ldg4(nb + 7);
    const Vec4 nx = r.pos & 1u ? lx : hx, fx = r.pos & 1u ? hx : lx;
    const Vec4 ny = r.pos & 2u ? ly : hy, fy = r.pos & 2u ? hy : ly;
    const Vec4 nz = r.pos & 4u ? lz : hz, fz = r.pos & 4u ? hz : lz;
    const float tEntry = h.t;
    float tn[4];
    bool hit[4];
#pragma unroll
    for (int j = 0; j < 4; j++) {
        const float xn = fmaf(comp(nx, j), r.ix, -odx), xf = fmaf(comp(fx, j), r.ix, -odx);
        const float yn = fmaf(comp(ny, j), r.iy, -ody), yf = fmaf(comp(fy, j), r.iy, -ody);
        const float zn = fmaf(comp(nz, j), r.iz, -odz), zf = fmaf(comp(fz, j), r.iz, -odz);
        tn[j] = fmaxf(fmaxf(xn, yn), fmaxf(zn, r.tmin));
        hit[j] = tn[j] <= fminf(fminf(xf, yf), fminf(zf, tEntry));
    }
    const int c[4] = { asInt(cf.x), asInt(cf.y), asInt(cf.z), asInt(cf.w) };
    const int k[4] = { asInt(kf.x), asInt(kf.y), asInt(kf.z), asInt(kf.w) };
#pragma unroll
    for (int j = 0; j < 4; j++)
        if (hit[j] && k[j] > 0)
            for (int t = c[j]; t < c[j] + k[j]; t++) if (triTest<TRI>(tri, t, r, h) && ANY) return true;
    constexpr float kInf = 3.4e38f;
    float ht[4];
    int hc[4];
#pragma unroll
    for (int j = 0; j < 4; j++) { ht[j] = hit[j] && k[j] == 0 && tn[j] <= h.t ? tn[j] : kInf; hc[j] = c[j]; }
    auto cswap = [&](int a, int b) {
        if (ht[b] < ht[a]) { const float tf = ht[a]; ht[a] = ht[b]; ht[b] = tf; const int ti = hc[a]; hc[a] = hc[b]; hc[b] = ti; }
    };
    cswap(0, 1); cswap(2, 3); cswap(0, 2); cswap(1, 3); cswap(1, 2);
    if (ht[0] == kInf) {
        if (st.sp == st.floor) return true;
        ts.node = st.pop();
        return false;
    }
    if (ht[3] < kInf) st.push(hc[3]);
    if (ht[2] < kInf) st.push(hc[2]);
    if (ht[1] < kInf) st.push(hc[1]);
    ts.node = hc[0];
    return false;
}

template <bool ANY, int TRI, class Stack>
__device__ __forceinline__ bool traceBVH4(const Vec4* __restrict__ nodes, const Vec4* __restrict__ tri, int root,
                                          const RayPre& r, Hit& h, Stack& st) {
    TravState ts; travBegin(ts, root);
    while (!stepBVH4<ANY, TRI>(nodes, tri, r, h, st, ts)) {}
    return h.tri >= 0;
}

// ------------------------------------------------------------------------------------------------ compressed BVH8
__device__ __forceinline__ uint32_t byteOf(uint32_t w, int s) { return (w >> (8 * (s & 3))) & 255u; }

// One node visit: the node's 8 child slabs, its hit leaves, then the choice of the next node. Returns true when the
// traversal is over (nothing left to visit, or for ANY a hit was found). leaf(k) handles leaf primitive k (the node's
// primitive base + the offset in its slot's metadata) and returns true to end an any-hit traversal: a triangle test
// for a BLAS, an instance's BLAS traversal for a TLAS.
template <bool ANY, class Stack, class Leaf>
__device__ __forceinline__ bool stepBVH8With(const Vec4* __restrict__ nodes, const RayPre& r, Hit& h, Stack& st,
                                             TravState& ts, Leaf&& leaf) {
    const Vec4* nb = nodes + (size_t)ts.node * kBvh8Stride;
    const Vec4 f0 = ldg4(nb), f1 = ldg4(nb + 1), f2 = ldg4(nb + 2), f3 = ldg4(nb + 3), f4 = ldg4(nb + 4);
    const uint32_t ew = asUint(f0.w);
    const float sx = asFloatU((ew & 255u) << 23) * r.ix;
    const float sy = asFloatU(((ew >> 8) & 255u) << 23) * r.iy;
    const float sz = asFloatU(((ew >> 16) & 255u) << 23) * r.iz;
    const float ox = (f0.x - r.o.x) * r.ix, oy = (f0.y - r.o.y) * r.iy, oz = (f0.z - r.o.z) * r.iz;
    const uint32_t imask = ew >> 24;
    const uint32_t meta[2] = { asUint(f1.z), asUint(f1.w) };
    const uint32_t w1 = asUint(f1.x), leafMask = w1 >> 24;   // slots holding leaves (empty slots: in neither mask)
    const int childBase = (int)(w1 & 0xFFFFFFu), triBase = asInt(f1.y);
    const uint32_t qlx[2] = { asUint(f2.x), asUint(f2.y) }, qly[2] = { asUint(f2.z), asUint(f2.w) };
    const uint32_t qlz[2] = { asUint(f3.x), asUint(f3.y) }, qhx[2] = { asUint(f3.z), asUint(f3.w) };
    const uint32_t qhy[2] = { asUint(f4.x), asUint(f4.y) }, qhz[2] = { asUint(f4.z), asUint(f4.w) };
    // Phase 1: the 8 child slabs against the hit distance at node entry. With the ray's direction signs the near
    // plane of each axis is known up front (lo for a positive direction, hi for a negative one), so each slot needs
    // no per-axis min/max; the result is bit-identical to min/max of the two planes.
    const uint32_t nx[2] = { r.pos & 1u ? qlx[0] : qhx[0], r.pos & 1u ? qlx[1] : qhx[1] };
    const uint32_t fx[2] = { r.pos & 1u ? qhx[0] : qlx[0], r.pos & 1u ? qhx[1] : qlx[1] };
    const uint32_t ny[2] = { r.pos & 2u ? qly[0] : qhy[0], r.pos & 2u ? qly[1] : qhy[1] };
    const uint32_t fy[2] = { r.pos & 2u ? qhy[0] : qly[0], r.pos & 2u ? qhy[1] : qly[1] };
    const uint32_t nz[2] = { r.pos & 4u ? qlz[0] : qhz[0], r.pos & 4u ? qlz[1] : qhz[1] };
    const uint32_t fz[2] = { r.pos & 4u ? qhz[0] : qlz[0], r.pos & 4u ? qhz[1] : qlz[1] };
    const float tEntry = h.t;
    uint32_t hitSlots = 0;
#pragma unroll
    for (int s = 0; s < 8; s++) {
        const int w = s >> 2;
        const float xn = fmaf((float)byteOf(nx[w], s), sx, ox), xf = fmaf((float)byteOf(fx[w], s), sx, ox);
        const float yn = fmaf((float)byteOf(ny[w], s), sy, oy), yf = fmaf((float)byteOf(fy[w], s), sy, oy);
        const float zn = fmaf((float)byteOf(nz[w], s), sz, oz), zf = fmaf((float)byteOf(fz[w], s), sz, oz);
        const float tn = fmaxf(fmaxf(xn, yn), fmaxf(zn, r.tmin));
        const float tf = fminf(fminf(xf, yf), fminf(zf, tEntry));
        hitSlots |= (uint32_t)(tn <= tf) << s;
    }
    // Phase 2: the triangles of the hit leaves as one bit set relative to the node's triangle base (a node's leaves
    // hold their triangles at increasing offsets in slot order, at most 24 in all), then one flat loop over its bits:
    // the wave iterates as often as its lane with the most triangles, not once per leaf and again per triangle.
    uint32_t lb = hitSlots & leafMask, tm = 0;
    while (lb) {
        const int s = __builtin_ctz(lb);
        lb &= lb - 1;
        const uint32_t m = byteOf(s < 4 ? meta[0] : meta[1], s);
        tm |= ((2u << ((m >> 5) & 3u)) - 1u) << (m & 31u);
    }
    while (tm) {
        const int j = __builtin_ctz(tm);
        tm &= tm - 1;
        if (leaf(triBase + j) && ANY) return true;
    }
    // internal hits in octant order: bit k = slot k ^ octant (permute by the three octant bits)
    uint32_t hits = hitSlots & imask;
    if (r.oct & 1u) hits = ((hits & 0x55u) << 1) | ((hits >> 1) & 0x55u);
    if (r.oct & 2u) hits = ((hits & 0x33u) << 2) | ((hits >> 2) & 0x33u);
    if (r.oct & 4u) hits = ((hits & 0x0Fu) << 4) | ((hits >> 4) & 0x0Fu);
    if (hits) {
        if (ts.gHits) st.push(makeGroup(ts.gBase, (ts.gMask << 8) | ts.gHits));
        ts.gBase = (uint32_t)childBase; ts.gMask = imask; ts.gHits = hits;
    }
    while (ts.gHits == 0) {
        if (st.sp == st.floor) return true;
        const Group e = st.pop();
        const uint32_t bits = (uint32_t)(e >> 32);
        ts.gBase = (uint32_t)e; ts.gMask = bits >> 8; ts.gHits = bits & 255u;
    }
    const uint32_t k = (uint32_t)__builtin_ctz(ts.gHits);
    ts.gHits &= ts.gHits - 1;
    const uint32_t slot = k ^ r.oct;
    ts.node = (int)ts.gBase + __builtin_popcount(ts.gMask & ((1u << slot) - 1u));
    return false;
}

template <bool ANY, int TRI, class Stack>
__device__ __forceinline__ bool stepBVH8(const Vec4* __restrict__ nodes, const Vec4* __restrict__ tri, const RayPre& r,
                                         Hit& h, Stack& st, TravState& ts) {
    return stepBVH8With<ANY>(nodes, r, h, st, ts, [&](int t) { return triTest<TRI>(tri, t, r, h); });
}

template <bool ANY, int TRI, class Stack>
__device__ __forceinline__ bool traceBVH8(const Vec4* __restrict__ nodes, const Vec4* __restrict__ tri, int root,
                                          const RayPre& r, Hit& h, Stack& st) {
    TravState ts; travBegin(ts, root);
    while (!stepBVH8<ANY, TRI>(nodes, tri, r, h, st, ts)) {}
    return h.tri >= 0;
}

}  // namespace cwp
