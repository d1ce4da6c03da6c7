// device/kernels.h - device code of the renderer: constants, path-state and scene records, shading, and the kernel
// templates per (BVH layout V, triangle test TRI, LDS stack SL). Each layout's kernels are instantiated in exactly one
// translation unit (kernels_v2.hip, kernels_v4.hip, kernels_v8.hip) through pickKernels<V>; renderer.hip holds the
// layout-independent kernels and the host driver.
//
// Design (MI355X-first, not the reference's launch structure):
//  * Persistent grids: every kernel launches a fixed grid of single-wave workgroups sized to the occupancy limit and
//    walks its queue in wave-sized chunks (grid stride). Queue lengths are read from device memory, so a frame is
//    enqueued once, with no count readback between bounces.
//  * Queue compaction per wave: the continuation and shadow decisions are ballots; one lane reserves the wave's block
//    of slots with a single atomic and every lane writes at its mbcnt rank (reference: per-thread atomicAdd appends,
//    src/render/wavefront.cu:119,135).
//  * Bounce 0 generates camera rays in registers (reference: separate k_raygen, src/render/wavefront.cu:42-58), so
//    camera rays never round-trip through HBM.
//  * Radiance is accumulated per path (pL) with plain read-modify-writes: each path has at most one shadow ray per
//    bounce, so no atomics are needed; a final kernel sums the K samples of each pixel in sample order. Results are
//    therefore independent of scheduling (the reference's float atomics into the pixel are not,
//    src/render/wavefront.cu:71,100).
//  * Traversal stacks live in LDS with spill to a per-thread global slab (device/traverse.h).
// Shading restates the reference's legacy shadeOne (src/render/wavefront.cu:94-136): sky on a miss, stop past maxB,
// geometric (and optional interpolated) normal faced to the ray, origin offset eps, sun next-event estimation by a
// uniform cone sample with contribution T*albedo*sunE*cos/pi, cosine-weighted continuation rejected below the
// geometric surface, throughput *= albedo.
#pragma once
#include "core/bsdf.h"
#include "core/sampling.h"
#include "device/renderer.h"
#include "device/traverse.h"
#include "host/texture.h"
#include <hip/hip_runtime.h>

namespace cwp {
namespace dev {

constexpr int kWave = 64;
constexpr int kBlock = 64;       // one wave per workgroup: LDS stacks are sized per wave
static_assert(kBlock == kStackStride, "the LDS stack is strided by the workgroup size");
constexpr int kParts = 8;        // queue partitions, one per XCD (workgroup w runs on XCD w % 8)
// Per-depth counters: one 128-byte line per partition, so the atomics of different partitions never meet on a line.
// Line x: [kCQueue] path-queue length of partition x, [kCShadow] shadow rays it produced, [kCMiss] its misses,
// [kCUnocc] its unoccluded shadow rays. Misses and unoccluded rays are summed per wave and added once at kernel end.
constexpr int kLine = 32;
constexpr int kCntStride = kParts * kLine;
constexpr int kCQueue = 0, kCShadow = 1, kCMiss = 2, kCUnocc = 3;
// rolling batches (steps instead of passes) also keep per partition and step: [kCIn] rays the step traces (the previous
// step's continuations + the step's camera rays), [kCMissCam] camera-ray misses
constexpr int kCIn = 4, kCMissCam = 5;
// straggler hand-off: [kCTail] traversals partition x's trace+shade waves handed to the tail kernel
constexpr int kCTail = 6, kCShadowTail = 7;   // ... and [kCShadowTail] the shadow kernel's
// [kCTail2] / [kCShadowTail2]: traversals the first tail kernels handed on to the second ones
constexpr int kCTail2 = 8, kCShadowTail2 = 9;

// Lights and materials of the full integrator (device pointers; null / 0 for the legacy integrator)
struct DevLights {
    const Material* mats;        // per material id
    const float* triLightPdf;    // per triangle (leaf position): light-selection probability / area, 0 if not emissive
    const int* emTri;            // emissive triangles (leaf positions), in the order of their power CDF
    const float* emCdf;          // nEm + 1 entries
    int nEm;
    const Vec4* envTex;          // environment radiance, row 0 = +z
    const float* envMarg;        // envH + 1
    const float* envCond;        // envH rows of envW + 1
    int envW, envH;
    float envRot;                // radians about +z
    float pSun, pTri, pEnv;      // light-type selection probabilities of next-event estimation
    int lightMode, rr, rrDepth, gdEnergy, envUniform, hasEnv;
    // two-level scenes: this frame's world-space emitters (3 float4 each: (v0, leaf bits), (e1, material bits),
    // (e2, 0)) and per material the light pdf per unit area (emTri / triLightPdf are unused)
    const Vec4* emTris;
    const float* matLightPdf;
    // textures: bindings per material (null: none), corner uvs / tangents per leaf position, images and their texels,
    // the sRGB decode table
    const MaterialMaps* maps;
    int nMaps, texNearest, bumpShadow;
    const Vec4* uvs;
    const Vec4* tans;
    const TexImage* texImages;
    const uint32_t* texels;
    const float* srgbTable;
};

// Fills L from the full-integrator data of S, uploading each array with up(pointer, vector) (host side). The sun is
// always light-sampled; emitters and the environment are light-sampled unless the light mode is BSDF-only; the light
// types that are present share the selection probability equally (reference: initFull,
// src/render/wavefront.cu:530-533).
template <class Up>
bool makeDevLights(const SceneUpload& S, Up&& up, DevLights& L, std::string& err) {
    L = DevLights{};
    if (!S.full) return true;
    if (!up(L.mats, S.mats, err)) return false;
    if (S.inst) {   // per-frame emitters: buffers of this scene's size, filled for each frame
        L.nEm = S.numEmInst;
        if (L.nEm > 0 && (!up(L.emTris, std::vector<Vec4>(3 * (size_t)L.nEm), err) ||
                          !up(L.emCdf, std::vector<float>((size_t)L.nEm + 1), err) ||
                          !up(L.matLightPdf, std::vector<float>(S.mats.size()), err)))
            return false;
    } else {
        if (!up(L.triLightPdf, S.triLightPdf, err) || !up(L.emTri, S.emTri, err) || !up(L.emCdf, S.emCdf, err)) return false;
        L.nEm = (int)S.emTri.size();
    }
    L.hasEnv = S.env && S.env->w > 0 ? 1 : 0;
    if (L.hasEnv) {
        if (!up(L.envTex, S.env->texels, err) || !up(L.envMarg, S.env->marg, err) || !up(L.envCond, S.env->cond, err)) return false;
        L.envW = S.env->w; L.envH = S.env->h;
        L.envRot = S.envRotDeg * 3.14159265f / 180.f;
    }
    const bool sun = S.sunOn != 0, tri = L.nEm > 0 && S.lightMode != kLightBSDF, env = L.hasEnv && S.lightMode != kLightBSDF;
    const int nt = (sun ? 1 : 0) + (tri ? 1 : 0) + (env ? 1 : 0);
    L.pSun = sun ? 1.f / (float)nt : 0.f; L.pTri = tri ? 1.f / (float)nt : 0.f; L.pEnv = env ? 1.f / (float)nt : 0.f;
    L.lightMode = S.lightMode; L.rr = S.rr; L.rrDepth = S.rrDepth; L.gdEnergy = S.gdEnergy; L.envUniform = S.envUniform;
    if (!S.maps.empty()) {
        if (!up(L.maps, S.maps, err) || !up(L.uvs, S.uvs, err) || !up(L.texImages, S.texImages, err) ||
            !up(L.texels, S.texels, err) || !up(L.srgbTable, std::vector<float>(srgbDecodeTable(), srgbDecodeTable() + 256), err))
            return false;
        if (!S.tans.empty() && !up(L.tans, S.tans, err)) return false;
        L.nMaps = (int)S.maps.size(); L.texNearest = S.texNearest; L.bumpShadow = S.bumpShadow;
    }
    return true;
}

struct DevScene {
    const Vec4* nodes;    // BVH layout (BVH8: 128-byte node stride)
    const Vec4* tri;      // traversal triangle data (Woop rows or corners)
    const Vec4* verts;    // corners (full-integrator shading, Moller-Trumbore and watertight tests)
    const Vec4* geo;      // per triangle: cross(e1, e2) (unnormalised geometric normal), material id bits
    const Vec4* nrm;      // per-corner normals or null
    const Vec4* albedo;   // per material
    int root;
    // two-level scenes: TLAS over instance ids and its root, instance records; nodes / tri hold every BLAS. The TLAS is
    // a BVH2 (64-byte nodes, leaves list instance ids) or, with BVH8 BLASes, a compressed BVH8 (128-byte device nodes,
    // root 0) whose leaf primitives k name instance tlasPrim[k]
    const Vec4* tlas;
    int tlasRoot;
    int tlasWide;
    const int* tlasPrim;
    const InstanceRecord* inst;
    DevLights Lt;         // full integrator
};

// Path state. Path id i = pixel * K + sample: the 64 lanes of a wave start as 64 samples of one pixel, and each
// partition's slice of the id space is a band of the image.
// The rays of depth d >= 1 are stored at their queue slot, not at their path id: set d & 1 holds them, the slot is the
// position the depth-(d-1) kernel appended them at, and the path id rides in O.w. A wave's appends are one contiguous
// block (base + mbcnt), so the next bounce reads its 64 rays as whole cache lines, with no id-queue load in front of
// them; only the radiance L stays indexed by path id.
struct DevPaths {
    Vec4 *O[2], *D[2], *T[2];  // continuation rays by queue slot: (origin, path id bits), (direction, BSDF pdf), throughput
    Vec4* L;                   // radiance per path id
    Vec4* H;                   // closest hit by queue slot (t, u, v, triangle bits) when tracing and shading are split
    Vec4 *sO, *sD, *sC;        // shadow rays of the current bounce: (origin, path id bits), (dir, tmax), (contribution)
    int* cnt;                  // kCntStride counters per depth (one line per partition)
    void* spill;               // traversal-stack spill slab, spillCap entries per thread
    int spillCap;
    // continuation sort (null when off): continuations are appended to the staging rays stO / stD / stT with their
    // direction bucket in qk, and k_qcount / k_qscatter move them into set (d + 1) & 1 grouped by bucket within each
    // partition; bcnt holds per depth and partition the bucket sizes and fill cursors (kSortWords ints per depth)
    Vec4 *stO = nullptr, *stD = nullptr, *stT = nullptr;
    uint8_t* qk = nullptr;
    int* bcnt = nullptr;
    // first-bounce sort (the pass's depth-0 launch only): each camera ray's continuation goes to the staging rays at
    // its own slot with its elevation bucket in qk (kNoKey: none), and k_qsort_first groups each tile of kFirstTile
    // slots, a pixel's samples, by bucket into the depth-1 queue
    int firstSort = 0;
    // straggler hand-off (tail null: off): a trace+shade wave whose tracing lanes drop below tailLanes stops tracing;
    // those lanes write their traversal (queue slot, hit so far, node group, stack) as a tail record and k_bounce_tail
    // finishes them in full waves and shades them. tailStride: Vec4s per record (3 + two stack entries per Vec4).
    Vec4* tail = nullptr;
    int tailStride = 0, tailLanes = 0;
    // second hand-off level (null: off): the tail kernels hand their own stragglers on, in the same way, to tail2
    Vec4* tail2 = nullptr;
};
constexpr uint8_t kNoKey = 255;

// First tail record of partition `part` whose input queue starts at `inStart`: each 64-entry chunk of a partition
// hands off at most lanes - 1 traversals, and floor((s + n) / 64) - floor(s / 64) + 1 >= ceil(n / 64), so the
// partitions' regions do not overlap; all of them fit in ((entries >> 6) + kParts + 1) * (lanes - 1) records.
__device__ __forceinline__ size_t tailFirst(size_t inStart, int part, int lanes) {
    return ((inStart >> 6) + (size_t)part) * (size_t)(lanes - 1);
}
// hit triangle of a ray whose traversal was handed to the tail kernel (not a miss: its shading waits for the tail)
constexpr int kHandedOff = -2;

// Continuation sort: 64 direction buckets, an 8 x 8 grid over the octahedral map of the unit sphere (upper hemisphere
// in the central square, the lower one folded onto the corners), so that a wave of one bucket shares a narrow cone of
// directions instead of the whole cosine lobe of one pixel's samples.
constexpr int kBuckets = 64;
constexpr int kSortWords = kParts * 2 * kBuckets;   // per depth: [partition][0: size, 1: cursor][bucket]
__device__ __forceinline__ int dirBucket(Vec3 d) {
    const float a = 1.f / (fabsf(d.x) + fabsf(d.y) + fabsf(d.z));
    float u = d.x * a, v = d.y * a;
    if (d.z < 0.f) {
        const float fu = copysignf(1.f - fabsf(v), u), fv = copysignf(1.f - fabsf(u), v);
        u = fu; v = fv;
    }
    // fmaxf first: a NaN component lands in bucket 0
    const int iu = (int)fminf(fmaxf((u + 1.f) * 4.f, 0.f), 7.f), iv = (int)fminf(fmaxf((v + 1.f) * 4.f, 0.f), 7.f);
    return iu + 8 * iv;
}

// Elevation bucket: the direction's z (the world is z-up) in 64 equal steps. Among the bounce rays leaving one point,
// those heading down into the geometry traverse much longer than those heading up and out.
__device__ __forceinline__ int elevBucket(Vec3 d) {
    return (int)fminf(fmaxf((d.z + 1.f) * 32.f, 0.f), 63.f);   // fmaxf first: NaN lands in bucket 0
}

// First-bounce sort: tiles of kFirstTile consecutive camera rays (one thread each; at 256 samples per pixel in flight,
// one pixel) have their continuations grouped by elevation bucket.
constexpr int kFirstTile = 256;
struct TileGroupLds { int h[kBuckets], first[kBuckets], base; };
// One tile's grouping, called by all kFirstTile threads of the workgroup (it holds barriers): key = this thread's
// bucket (kNoKey: no continuation). Counts per bucket in LDS (a thread's rank in its bucket from the LDS atomic),
// an exclusive prefix over the buckets by the first wave, whose last lane reserves the tile's total in `*queue` with
// one atomic. Returns the continuation's queue position relative to the partition start (-1 for kNoKey).
__device__ __forceinline__ int tileGroup(int key, int* queue, TileGroupLds& g) {
    const int t = (int)threadIdx.x;
    if (t < kBuckets) g.h[t] = 0;
    __syncthreads();
    int rank = 0;
    if (key != kNoKey) rank = atomicAdd(&g.h[key & (kBuckets - 1)], 1);
    __syncthreads();
    if (t < kBuckets) {   // exclusive prefix over the first wave's 64 lanes; the last lane reserves the total
        const int c = g.h[t];
        int s = c;
#pragma unroll
        for (int o = 1; o < kBuckets; o <<= 1) {
            const int y = __shfl_up(s, o, kBuckets);
            if (t >= o) s += y;
        }
        g.first[t] = s - c;
        if (t == kBuckets - 1) g.base = s > 0 ? atomicAdd(queue, s) : 0;
    }
    __syncthreads();
    const int pos = key != kNoKey ? g.base + g.first[key & (kBuckets - 1)] + rank : -1;
    __syncthreads();   // g is reused by the next tile
    return pos;
}

template <int V> struct StackEntry;
template <> struct StackEntry<2> { using type = int; };
template <> struct StackEntry<4> { using type = int; };
template <> struct StackEntry<8> { using type = Group; };

template <int V, bool ANY, int TRI, class St>
__device__ __forceinline__ bool traceScene(const DevScene& S, const RayPre& r, Hit& h, St& st) {
    if constexpr (V == 2) return traceBVH2<ANY, TRI>(S.nodes, S.tri, S.root, r, h, st);
    else if constexpr (V == 4) return traceBVH4<ANY, TRI>(S.nodes, S.tri, S.root, r, h, st);
    else return traceBVH8<ANY, TRI>(S.nodes, S.tri, S.root, r, h, st);
}

// instance rows applied to a point / a vector (sums left to right), and the normal transform inv^T n
__device__ __forceinline__ Vec3 xfRowsP(const Vec4* r, Vec3 p) {
    return v3(r[0].x * p.x + r[0].y * p.y + r[0].z * p.z + r[0].w, r[1].x * p.x + r[1].y * p.y + r[1].z * p.z + r[1].w,
              r[2].x * p.x + r[2].y * p.y + r[2].z * p.z + r[2].w);
}
__device__ __forceinline__ Vec3 xfRowsV(const Vec4* r, Vec3 p) {
    return v3(r[0].x * p.x + r[0].y * p.y + r[0].z * p.z, r[1].x * p.x + r[1].y * p.y + r[1].z * p.z,
              r[2].x * p.x + r[2].y * p.y + r[2].z * p.z);
}
__device__ __forceinline__ Vec3 xfRowsN(const Vec4* r, Vec3 n) {
    return v3(r[0].x * n.x + r[1].x * n.y + r[2].x * n.z, r[0].y * n.x + r[1].y * n.y + r[2].y * n.z,
              r[0].z * n.x + r[1].z * n.y + r[2].z * n.z);
}

// The BLAS of instance k for a world ray (reference: travInst's instance leaf, src/accel/traverse_inst.cuh:26-29): the
// ray in object space (o' = inv o, d' = inv d, not normalised, so hit distances are shared with the world ray); the
// BLAS's stack entries sit above the TLAS's (st.floor). An instance that lowers the hit distance becomes the hit
// instance.
template <int V, bool ANY, int TRI, class St>
__device__ __forceinline__ void enterInstance(const DevScene& S, int k, Vec3 o, Vec3 d, float tmin, Hit& h, St& st) {
    const InstanceRecord* I = S.inst + k;
    const Vec4 r[3] = { I->inv[0], I->inv[1], I->inv[2] };
    const RayPre rb = prepareRay<TRI>(xfRowsP(r, o), xfRowsV(r, d), tmin);
    const float before = h.t;
    const int floor = st.floor;
    st.floor = st.sp;
    TravState ts; travBegin(ts, I->root);
    if constexpr (V == 2) { while (!stepBVH2<ANY, TRI>(S.nodes, S.tri, rb, h, st, ts)) {} }
    else if constexpr (V == 4) { while (!stepBVH4<ANY, TRI>(S.nodes, S.tri, rb, h, st, ts)) {} }
    else { while (!stepBVH8<ANY, TRI>(S.nodes, S.tri, rb, h, st, ts)) {} }
    st.floor = floor;
    if (h.t < before) h.inst = k;
}

// Two-level traversal over a compressed BVH8 TLAS: the BLAS node step with instances as leaf primitives, each hit
// leaf's instances entered in slot order; the TLAS's groups share the lane stack with the BLAS traversal above them.
// The closest hit is the BVH2 TLAS's (both visit every instance whose box the ray enters before the hit).
template <bool ANY, int TRI, class St>
__device__ __forceinline__ bool traceInstWide(const DevScene& S, Vec3 o, Vec3 d, float tmin, Hit& h, St& st) {
    const RayPre r = prepareRay<kTriMT>(o, d, tmin);
    const int* __restrict__ prim = S.tlasPrim;
    TravState ts; travBegin(ts, 0);
    while (!stepBVH8With<ANY>(S.tlas, r, h, st, ts, [&](int k) {
        enterInstance<8, ANY, TRI>(S, prim[k], o, d, tmin, h, st);
        return h.tri >= 0;
    })) {}
    return h.tri >= 0;
}

// Two-level traversal (reference: travInst, src/accel/traverse_inst.cuh:10-41): the TLAS is a BVH2 whose leaves list
// instance ids (or, for BVH8 BLASes, a compressed BVH8: traceInstWide); the TLAS's stack entries share the lane stack
// with the BLAS traversal above them.
template <int V, bool ANY, int TRI, class St>
__device__ __forceinline__ bool traceInst(const DevScene& S, Vec3 o, Vec3 d, float tmin, Hit& h, St& st) {
    using E = typename StackEntry<V>::type;
    if constexpr (V == 8)
        if (S.tlasWide) return traceInstWide<ANY, TRI>(S, o, d, tmin, h, st);
    const RayPre r = prepareRay<kTriMT>(o, d, tmin);
    const float odx = r.o.x * r.ix, ody = r.o.y * r.iy, odz = r.o.z * r.iz;
    int node = S.tlasRoot;
    for (;;) {
        const Vec4* nb = S.tlas + (size_t)node * 4;
        const Vec4 n0 = ldg4(nb), n1 = ldg4(nb + 1), nz = ldg4(nb + 2), mt = ldg4(nb + 3);
        int c0 = asInt(mt.x), c1 = asInt(mt.y);
        const int k0 = asInt(mt.z), k1 = asInt(mt.w);
        const float a0 = fmaf(n0.x, r.ix, -odx), a1 = fmaf(n0.y, r.ix, -odx), a2 = fmaf(n0.z, r.iy, -ody), a3 = fmaf(n0.w, r.iy, -ody);
        const float a4 = fmaf(nz.x, r.iz, -odz), a5 = fmaf(nz.y, r.iz, -odz);
        const float b0 = fmaf(n1.x, r.ix, -odx), b1 = fmaf(n1.y, r.ix, -odx), b2 = fmaf(n1.z, r.iy, -ody), b3 = fmaf(n1.w, r.iy, -ody);
        const float b4 = fmaf(nz.z, r.iz, -odz), b5 = fmaf(nz.w, r.iz, -odz);
        const float t0n = fmaxf(fmaxf(fminf(a0, a1), fminf(a2, a3)), fmaxf(fminf(a4, a5), r.tmin));
        const float t0f = fminf(fminf(fmaxf(a0, a1), fmaxf(a2, a3)), fminf(fmaxf(a4, a5), h.t));
        const float t1n = fmaxf(fmaxf(fminf(b0, b1), fminf(b2, b3)), fmaxf(fminf(b4, b5), r.tmin));
        const float t1f = fminf(fminf(fmaxf(b0, b1), fmaxf(b2, b3)), fminf(fmaxf(b4, b5), h.t));
        bool h0 = t0n <= t0f, h1 = t1n <= t1f;
        if (h0 && k0 > 0) {
            for (int t = c0; t < c0 + k0; t++) {
                enterInstance<V, ANY, TRI>(S, t, o, d, tmin, h, st);
                if (ANY && h.tri >= 0) return true;
            }
            h0 = false;
        }
        if (h1 && k1 > 0) {
            for (int t = c1; t < c1 + k1; t++) {
                enterInstance<V, ANY, TRI>(S, t, o, d, tmin, h, st);
                if (ANY && h.tri >= 0) return true;
            }
            h1 = false;
        }
        if (!h0 && !h1) {
            if (st.sp == 0) break;
            node = (int)st.pop();
        } else if (h0 && h1) {
            if (t1n < t0n) { const int x = c0; c0 = c1; c1 = x; }
            st.push((E)(uint32_t)c1);
            node = c0;
        } else node = h0 ? c0 : c1;
    }
    return h.tri >= 0;
}

// closest / any hit of one ray in a single- or two-level scene
template <int V, bool ANY, int TRI, bool INST, class St>
__device__ __forceinline__ bool traceRay(const DevScene& S, Vec3 o, Vec3 d, float tmin, Hit& h, St& st) {
    st.sp = 0;
    if constexpr (INST) return traceInst<V, ANY, TRI>(S, o, d, tmin, h, st);
    else return traceScene<V, ANY, TRI>(S, prepareRay<TRI>(o, d, tmin), h, st);
}

__device__ __forceinline__ int laneId() { return (int)__lane_id(); }

// Reserves popcount(mask) consecutive slots of *counter for the lanes in mask (one atomic per wave) and returns this
// lane's slot. Must be called by the whole wave with a wave-uniform, non-zero mask.
__device__ __forceinline__ int waveAppend(uint64_t mask, int* counter) {
    const int leader = __builtin_ctzll(mask);
    int base = 0;
    if (laneId() == leader) base = atomicAdd(counter, (int)__builtin_popcountll(mask));
    base = __builtin_amdgcn_readlane(base, leader);
    const int rank = (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
    return base + rank;
}

// adds a wave-uniform count to *counter once per wave (nothing when it is zero)
__device__ __forceinline__ void waveAdd(int count, int* counter) {
    if (count && laneId() == 0) atomicAdd(counter, count);
}

// counter `which` of partition `part` in a depth's counter block
__device__ __forceinline__ int* ctr(int* cnt, int part, int which) { return cnt + part * kLine + which; }
__device__ __forceinline__ int ctrv(const int* cnt, int part, int which) { return cnt[part * kLine + which]; }

// start of partition `part` in a region laid out as the partitions' counts `which` of a counter block (default: the
// queue lengths of a depth)
__device__ __forceinline__ int partStart(const int* cnt, int part, int which = kCQueue) {
    int s = 0;
    for (int y = 0; y < part; y++) s += ctrv(cnt, y, which);
    return s;
}

// camera ray of path (pix, smp): filter importance sampling (Blackman-Harris table or 1 px box), pinhole or thin lens
// (reference: src/render/wavefront.cu:43-57)
__device__ __forceinline__ void cameraRay(const FrameParams& F, int pix, uint32_t smp, Vec3& o, Vec3& d) {
    const Camera& c = F.cam;
    const int x = pix % c.w, y = pix / c.w;
    const float u0 = rnd((uint32_t)pix, smp, 0, 0, F.seed), u1 = rnd((uint32_t)pix, smp, 0, 1, F.seed);
    float jx, jy;
    if (F.filterBH) { jx = tableLookup(F.ftab, kFilterTable, u0); jy = tableLookup(F.ftab, kFilterTable, u1); }
    else { jx = u0 - 0.5f; jy = u1 - 0.5f; }
    const float dx = ((float)x + 0.5f + jx - c.px) / c.fx, dy = -((float)y + 0.5f + jy - c.py) / c.fy;
    const Vec3 dir = c.ax * dx + c.ay * dy - c.az;
    d = normalize(dir);
    o = c.pos;
    if (c.lensR > 0.f) {
        float lx, ly;
        concentricDisk(rnd((uint32_t)pix, smp, 0, 2, F.seed), rnd((uint32_t)pix, smp, 0, 3, F.seed), lx, ly);
        Vec3 pf = c.pos + dir * c.focusD;
        o = c.pos + c.ax * (lx * c.lensR) + c.ay * (ly * c.lensR);
        d = normalize(pf - o);
    }
}

// ------------------------------------------------------------------------------------------------------- kernels
// Legacy shading of one closest hit of path i (reference: shadeOne, src/render/wavefront.cu:94-136). Writes the
// path's radiance on a miss (depth 0 also initialises it on a hit) and returns the continuation (cont: po, nd, Tn) and
// the shadow ray (shadow: po, sdir, scon) for the caller's wave-level appends.
// Full-integrator shading also returns the shadow ray's own origin (so) and extent (stmax) and the continuation's
// BSDF pdf (stored in D.w for the MIS weight of its hit).
struct ShadeOut { bool cont, shadow, miss; Vec3 po, nd, Tn, sdir, scon, so; float stmax, pdf; };

template <bool INST = false>
__device__ __forceinline__ ShadeOut shadeLegacy(const DevScene& S, const DevPaths& P, const FrameParams& F, int depth,
                                                int i, int pix, uint32_t smp, Vec3 o, Vec3 d, Vec3 T, const Hit& h) {
    ShadeOut out;
    out.cont = false; out.shadow = false; out.miss = false;
    out.po = v3(0.f, 0.f, 0.f); out.nd = out.po; out.Tn = out.po; out.sdir = out.po; out.scon = out.po; out.so = out.po;
    out.stmax = 1e30f; out.pdf = 0.f;
    if (h.tri < 0) {
        if (h.tri == kHandedOff) return out;
        out.miss = true;
        Vec3 c = T * F.sky;
        if (depth == 0) P.L[i] = v4(c, 0.f);
        else { Vec4 l = P.L[i]; l.x += c.x; l.y += c.y; l.z += c.z; P.L[i] = l; }
        return out;
    }
    if (depth == 0) P.L[i] = v4(0.f, 0.f, 0.f, 0.f);
    if (depth > F.maxB) return out;
    const Vec3 p = o + d * h.t;
    // one 16-byte record per triangle instead of its 48 bytes of corners: cross(e1, e2) is computed once per scene by
    // k_tri_geo with the operations this shader would apply, so the normal is bit-identical
    const Vec4 g = S.geo[h.tri];
    Vec4 xr[3];   // two-level scenes: object-space normals go to world space through inv^T
    if constexpr (INST) { const InstanceRecord* I = S.inst + h.inst; xr[0] = I->inv[0]; xr[1] = I->inv[1]; xr[2] = I->inv[2]; }
    Vec3 ng = INST ? normalize(xfRowsN(xr, xyz(g))) : normalize(xyz(g)), ns = ng;
    if (F.hasN) {
        Vec3 n0 = xyz(S.nrm[3 * h.tri]), n1 = xyz(S.nrm[3 * h.tri + 1]), n2 = xyz(S.nrm[3 * h.tri + 2]);
        Vec3 nn = n0 * (1.f - h.u - h.v) + n1 * h.u + n2 * h.v;
        if constexpr (INST) nn = xfRowsN(xr, nn);
        float l = length(nn);
        if (l > 0.f) ns = nn * (1.f / l);
        if (dot(ns, ng) < 0.f) ns = -ns;
    }
    if (dot(ng, d) > 0.f) { ng = -ng; ns = -ns; }
    const Vec3 alb = xyz(S.albedo[asInt(g.w)]);
    out.po = p + ng * F.eps;
    out.so = out.po;
    const uint32_t dd = (uint32_t)depth + 1u;
    const Vec3 l = uniformCone(F.sunTo, F.sunCos, rnd((uint32_t)pix, smp, dd, 2, F.seed), rnd((uint32_t)pix, smp, dd, 3, F.seed));
    const float cl = dot(ns, l);
    if (cl > 0.f && dot(ng, l) > 0.f) { out.shadow = true; out.sdir = l; out.scon = T * alb * F.sunE * (cl * kInvPi); }
    out.nd = cosineHemisphere(ns, rnd((uint32_t)pix, smp, dd, 0, F.seed), rnd((uint32_t)pix, smp, dd, 1, F.seed));
    if (dot(out.nd, ng) > 0.f) { out.cont = true; out.Tn = T * alb; }
    return out;
}

#include "device/shade_full.h"

// Appends of the shading results: continuations (with their next depth in T.w) to ray set `outSet` at outStart +
// a slot from *contCtr (or to the sort's staging rays), shadow rays at outStart + a slot from *shadowCtr. Compaction
// granularity: the wave (LANE false: one ballot and one atomic per wave, slots base + mbcnt) or the lane (LANE true:
// one atomic per appending lane, the reference's scheme). Returns the wave's number of misses.
template <bool LANE = false>
__device__ __forceinline__ int appendTo(const DevPaths& P, int* contCtr, int* shadowCtr, int outStart, int outSet,
                                        int nextDepth, int i, const ShadeOut& r) {
    if constexpr (LANE) {
        if (r.cont) {
            const int pos = outStart + atomicAdd(contCtr, 1);
            P.O[outSet][pos] = v4(r.po, asFloat(i)); P.D[outSet][pos] = v4(r.nd, r.pdf); P.T[outSet][pos] = v4(r.Tn, asFloat(nextDepth));
        }
        if (r.shadow) {
            const int pos = outStart + atomicAdd(shadowCtr, 1);
            P.sO[pos] = v4(r.so, asFloat(i)); P.sD[pos] = v4(r.sdir, r.stmax); P.sC[pos] = v4(r.scon, 0.f);
        }
        return (int)__builtin_popcountll(__ballot(r.miss));
    }
    const uint64_t mc = __ballot(r.cont), ms = __ballot(r.shadow), mm = __ballot(r.miss);
    if (mc) {
        const int pos = outStart + waveAppend(mc, contCtr);
        if (r.cont) {
            const Vec4 o = v4(r.po, asFloat(i)), d = v4(r.nd, r.pdf), t = v4(r.Tn, asFloat(nextDepth));
            if (P.qk) { P.stO[pos] = o; P.stD[pos] = d; P.stT[pos] = t; P.qk[pos] = (uint8_t)dirBucket(r.nd); }
            else { P.O[outSet][pos] = o; P.D[outSet][pos] = d; P.T[outSet][pos] = t; }
        }
    }
    if (ms) {
        const int pos = outStart + waveAppend(ms, shadowCtr);
        if (r.shadow) { P.sO[pos] = v4(r.so, asFloat(i)); P.sD[pos] = v4(r.sdir, r.stmax); P.sC[pos] = v4(r.scon, 0.f); }
    }
    return (int)__builtin_popcountll(mm);
}
// the appends of a depth-ordered pass: continuations to partition `part` of the depth + 1 queue (ray set
// (depth + 1) & 1), shadow rays to partition `part` of the depth's shadow queue
__device__ __forceinline__ int appendResults(const DevPaths& P, int* cnt, int part, int outStart, int depth, int i,
                                             const ShadeOut& r) {
    return appendTo(P, ctr(cnt + kCntStride, part, kCQueue), ctr(cnt, part, kCShadow), outStart, (depth + 1) & 1, depth + 1, i, r);
}

// The ray at queue slot `slot` of this depth and its path id i: at depth 0 the slot is the path id and the camera ray
// is computed from it; deeper, the stored ray of set depth & 1.
__device__ __forceinline__ void pathRay(const DevPaths& P, const FrameParams& F, int depth, int slot, int smpBase, int K,
                                        int& i, Vec3& o, Vec3& d, float& tmin) {
    if (depth == 0) { i = slot; cameraRay(F, i / K, (uint32_t)(smpBase + i % K), o, d); tmin = F.nearT; }
    else {
        const Vec4 O = P.O[depth & 1][slot];
        i = asInt(O.w); o = xyz(O); d = xyz(P.D[depth & 1][slot]); tmin = 0.f;
    }
}
// throughput and last BSDF pdf of the ray at queue slot `slot` (camera rays: 1 and 0)
__device__ __forceinline__ Vec3 pathThroughput(const DevPaths& P, int depth, int slot) {
    return depth == 0 ? v3(1.f, 1.f, 1.f) : xyz(P.T[depth & 1][slot]);
}
__device__ __forceinline__ float pathLastPdf(const DevPaths& P, int depth, int slot) {
    return depth == 0 ? 0.f : P.D[depth & 1][slot].w;
}

// Queue position of the f-th entry a wave hands out: the wave walks the chunks lw, lw + nw, lw + 2 nw, ... of its
// partition, 64 entries each (increasing in f).
__device__ __forceinline__ int fetchPos(uint32_t f, int lw, int nw) { return (lw + (int)(f >> 6) * nw) * kWave + (int)(f & 63u); }

// idle lanes of a persistent wave are refilled once at least this many are idle (kWave: only when all are, which is
// the plain one-ray-per-lane schedule)
constexpr int kRefill = 16;
// fused persistent bounce: the done lanes shade together once this many are done
constexpr int kShadeBatch = 32;

// Closest (ANY false) or any hit of one ray by BVH8 node steps, handed off once fewer than P.tailLanes lanes of the
// wave are still tracing: those lanes write their traversal as a tail record appended (through tailCtr) to the region
// of buf starting at record regionFirst (queue slot `slot`, the hit so far, the node group and the stack: (slot, t, u,
// v), (tri, node, group base, group mask), (group hits, stack depth, 0, 0), then two stack entries per Vec4) and mark
// the hit kHandedOff.
// One node step of a resumable traversal: the scene's BVH8 (INST false) or, for two-level scenes, the compressed BVH8
// TLAS, whose leaf primitives are instances entered (each BLAS traversed to the end) within the step; the world ray
// (o, d, tmin) is needed for the instances. True when the traversal is over.
template <bool ANY, int TRI, bool INST, class St>
__device__ __forceinline__ bool resumableStep(const DevScene& S, Vec3 o, Vec3 d, float tmin, const RayPre& r, Hit& h,
                                              St& st, TravState& ts) {
    if constexpr (INST) {
        const int* prim = S.tlasPrim;
        return stepBVH8With<ANY>(S.tlas, r, h, st, ts, [&](int k) {
            enterInstance<8, ANY, TRI>(S, prim[k], o, d, tmin, h, st);
            return h.tri >= 0;
        });
    } else {
        return stepBVH8<ANY, TRI>(S.nodes, S.tri, r, h, st, ts);
    }
}
// the ray of a resumable traversal: the world ray for the TLAS (the plain slab test's precomputation) or the BLAS ray
template <int TRI, bool INST>
__device__ __forceinline__ RayPre resumableRay(Vec3 o, Vec3 d, float tmin) {
    return INST ? prepareRay<kTriMT>(o, d, tmin) : prepareRay<TRI>(o, d, tmin);
}

template <bool ANY, int TRI, bool INST, class St>
__device__ __forceinline__ void stepOrHandOff(const DevScene& S, const DevPaths& P, Vec4* buf, size_t regionFirst,
                                              int* tailCtr, int slot, Vec3 o, Vec3 d, float tmin, const RayPre& r,
                                              Hit& h, St& st, TravState& ts) {
    for (;;) {
        if (resumableStep<ANY, TRI, INST>(S, o, d, tmin, r, h, st, ts)) return;
        if ((int)__builtin_popcountll(__ballot(1)) < P.tailLanes) break;   // the lanes still tracing
    }
    const size_t first = regionFirst + (size_t)waveAppend(__ballot(1), tailCtr);
    Vec4* rec = buf + first * (size_t)P.tailStride;
    rec[0] = v4(asFloat(slot), h.t, h.u, h.v);
    rec[1] = v4(asFloat(h.tri), asFloat(ts.node), asFloatU(ts.gBase), asFloatU(ts.gMask));
    rec[2] = v4(asFloatU(ts.gHits), asFloat(st.sp), asFloat(h.inst), 0.f);
    for (int e = 0; e < st.sp; e += 2) {
        const Group a = st.at(e), b = e + 1 < st.sp ? st.at(e + 1) : (Group)0;
        rec[3 + e / 2] = v4(asFloatU((uint32_t)a), asFloatU((uint32_t)(a >> 32)), asFloatU((uint32_t)b),
                            asFloatU((uint32_t)(b >> 32)));
    }
    h.tri = kHandedOff;
}
// The same from the root for a new ray; its records go to region `regionFirst` of buf.
template <bool ANY, int TRI, bool INST, class St>
__device__ __forceinline__ void traceOrHandOff(const DevScene& S, const DevPaths& P, Vec4* buf, size_t regionFirst,
                                               int* tailCtr, int slot, Vec3 o, Vec3 d, float tmin, Hit& h, St& st) {
    const RayPre r = resumableRay<TRI, INST>(o, d, tmin);
    TravState ts; travBegin(ts, INST ? 0 : S.root);
    st.sp = 0;
    stepOrHandOff<ANY, TRI, INST>(S, P, buf, regionFirst, tailCtr, slot, o, d, tmin, r, h, st, ts);
}

// A tail record back into a hit, a traversal state and the lane's stack; returns the record's queue slot.
template <class St>
__device__ __forceinline__ int resumeRecord(const Vec4* rec, Hit& h, TravState& ts, St& st) {
    const Vec4 a = rec[0], b = rec[1], c = rec[2];
    h.t = a.y; h.u = a.z; h.v = a.w; h.tri = asInt(b.x); h.inst = asInt(c.z);
    ts.node = asInt(b.y); ts.gBase = asUint(b.z); ts.gMask = asUint(b.w); ts.gHits = asUint(c.x);
    const int sp = asInt(c.y);
    st.sp = 0;
    for (int e = 0; e < sp; e += 2) {
        const Vec4 q = rec[3 + e / 2];
        st.push(((Group)asUint(q.y) << 32) | asUint(q.x));
        if (e + 1 < sp) st.push(((Group)asUint(q.w) << 32) | asUint(q.z));
    }
    return asInt(a.x);
}

// Shading of the traced ray at queue slot `slot` (path i); in the first-bounce sort's depth-0 launch its continuation
// is staged at its own slot for k_qsort_first instead of being appended.
template <bool INST, bool FULL>
__device__ __forceinline__ ShadeOut shadeSlot(const DevScene& S, const DevPaths& P, const FrameParams& F, int depth,
                                              int slot, int i, int smpBase, int K, Vec3 o, Vec3 d, const Hit& h) {
    // throughput is fetched after the traversal so that no registers hold it across the traversal loop
    const Vec3 T = pathThroughput(P, depth, slot);
    ShadeOut res;
    if constexpr (FULL) {
        const float lastPdf = pathLastPdf(P, depth, slot);
        res = shadeFull<INST>(S, P, F, depth, i, i / K, (uint32_t)(smpBase + i % K), o, d, T, lastPdf, h);
    } else {
        res = shadeLegacy<INST>(S, P, F, depth, i, i / K, (uint32_t)(smpBase + i % K), o, d, T, h);
    }
    if (P.firstSort) {   // depth 0: the continuation waits at its own slot for k_qsort_first
        P.qk[slot] = res.cont ? (uint8_t)elevBucket(res.nd) : kNoKey;
        if (res.cont) {
            P.stO[slot] = v4(res.po, asFloat(i)); P.stD[slot] = v4(res.nd, res.pdf);
            P.stT[slot] = v4(res.Tn, asFloat(1));
        }
        res.cont = false;
    }
    return res;
}

// One bounce of the legacy integrator over the queue of depth `depth` (depth 0: every path id, camera rays made here),
// traversal and shading fused. The queue is split into kParts partitions, one per XCD; the waves of XCD x walk
// partition x, and append their continuations to partition x of the next queue and their shadow rays to partition x of
// the shadow queue, so a path stays on one XCD (and in its L2) for its whole life. Partition x of the depth-(d+1) queue
// starts where the depth-d partitions before x end, which bounds it by the rays partition x traced.
// WAVES: minimum waves per SIMD the register allocation must allow (1 = unconstrained).
// TAIL (BVH8, single-level scenes): straggler hand-off. After each node visit a wave counts its lanes still tracing;
// once fewer than P.tailLanes are, they stop and write their traversal state as tail records (k_bounce_tail resumes
// them), and the wave shades its finished lanes and takes its next rays. The traversal resumes exactly where it
// stopped (same hit, node group and stack), so hits and images do not change.
template <int V, int TRI, int SL, int WAVES, bool INST, bool FULL = false, bool TAIL = false>
__global__ void __launch_bounds__(kBlock, WAVES) k_bounce(DevScene S, DevPaths P, const FrameParams* __restrict__ Fp, int depth,
                                                          int smpBase, int K) {
    static_assert(!TAIL || V == 8, "the hand-off resumes BVH8 traversals (two-level scenes: a BVH8 TLAS)");
    const FrameParams& F = *Fp;
    using E = typename StackEntry<V>::type;
    __shared__ E lds[SL > 0 ? SL * kBlock : 1];
    const int lane = laneId();
    const int gtid = blockIdx.x * kBlock + threadIdx.x;
    LaneStack<E, SL> st{ (LdsPtr<E>)(lds + threadIdx.x), (E*)P.spill + (size_t)gtid * P.spillCap, 0, 0 };
    const int part = (int)(blockIdx.x % kParts), lw = (int)(blockIdx.x / kParts), nw = (int)(gridDim.x / kParts);
    int* cnt = P.cnt + depth * kCntStride;
    const int n = ctrv(cnt, part, kCQueue);
    const int inStart = depth == 0 ? partStart(cnt, part) : partStart(cnt - kCntStride, part);
    const int outStart = partStart(cnt, part);
    int misses = 0;
    for (int base = lw * kWave; base < n; base += nw * kWave) {
        const int k = base + lane;
        int i = 0;
        ShadeOut res;
        res.cont = false; res.shadow = false; res.miss = false;
        if (k < n) {
            const int slot = inStart + k;
            Vec3 o, d; float tmin;
            pathRay(P, F, depth, slot, smpBase, K, i, o, d, tmin);
            Hit h; h.t = 1e30f; h.u = 0.f; h.v = 0.f; h.tri = -1; h.inst = -1;
            if constexpr (TAIL) {
                traceOrHandOff<false, TRI, INST>(S, P, P.tail, tailFirst(inStart, part, P.tailLanes), ctr(cnt, part, kCTail),
                                                 slot, o, d, tmin, h, st);
                res = shadeSlot<INST, FULL>(S, P, F, depth, slot, i, smpBase, K, o, d, h);   // nothing if handed off
            } else {
                traceRay<V, false, TRI, INST>(S, o, d, tmin, h, st);
                res = shadeSlot<INST, FULL>(S, P, F, depth, slot, i, smpBase, K, o, d, h);
            }
        }
        misses += appendResults(P, cnt, part, outStart, depth, i, res);
    }
    waveAdd(misses, ctr(cnt, part, kCMiss));
}

// Depth 0 with the first-bounce sort fused in: workgroups of kFirstTile threads (four waves, each with its own LDS
// stacks) trace and shade one tile of consecutive camera rays at a time, append the shadow rays per wave, and write
// the continuations straight into the depth-1 queue grouped by elevation bucket (tileGroup), instead of staging them
// for k_qsort_first. No hand-off: camera rays of one pixel need nearly the same traversals.
template <int V, int TRI, int SL, int WAVES, bool FULL>
__global__ void __launch_bounds__(kFirstTile, WAVES) k_bounce_first(DevScene S, DevPaths P, const FrameParams* __restrict__ Fp,
                                                                    int smpBase, int K) {
    const FrameParams& F = *Fp;
    using E = typename StackEntry<V>::type;
    __shared__ E lds[SL > 0 ? SL * kFirstTile : 1];
    __shared__ TileGroupLds g;
    const int lane = laneId();
    const int gtid = blockIdx.x * kFirstTile + threadIdx.x;
    LaneStack<E, SL> st{ (LdsPtr<E>)(lds + (threadIdx.x / kWave) * (SL > 0 ? SL : 1) * kWave + lane),
                         (E*)P.spill + (size_t)gtid * P.spillCap, 0, 0 };
    const int part = (int)(blockIdx.x % kParts), lb = (int)(blockIdx.x / kParts), nb = (int)(gridDim.x / kParts);
    int* cnt = P.cnt;
    const int n = ctrv(cnt, part, kCQueue);
    const int start = partStart(cnt, part);   // the depth-0 slots and the depth-1 queue of partition x start here
    int* queue = ctr(cnt + kCntStride, part, kCQueue);
    int misses = 0;
    for (int t0 = lb * kFirstTile; t0 < n; t0 += nb * kFirstTile) {   // block-uniform bound: every thread meets the barriers
        const int k = t0 + (int)threadIdx.x;
        int i = 0;
        ShadeOut res;
        res.cont = false; res.shadow = false; res.miss = false;
        if (k < n) {
            const int slot = start + k;
            Vec3 o, d; float tmin;
            pathRay(P, F, 0, slot, smpBase, K, i, o, d, tmin);
            Hit h; h.t = 1e30f; h.u = 0.f; h.v = 0.f; h.tri = -1; h.inst = -1;
            traceRay<V, false, TRI, false>(S, o, d, tmin, h, st);
            res = shadeSlot<false, FULL>(S, P, F, 0, slot, i, smpBase, K, o, d, h);
        }
        const bool cont = res.cont;
        res.cont = false;   // the per-wave appends take the shadow ray and the miss; the tile places the continuation
        misses += appendResults(P, cnt, part, start, 0, i, res);
        const int to = tileGroup(cont ? elevBucket(res.nd) : (int)kNoKey, queue, g);
        if (to >= 0) {
            P.O[1][start + to] = v4(res.po, asFloat(i)); P.D[1][start + to] = v4(res.nd, res.pdf);
            P.T[1][start + to] = v4(res.Tn, asFloat(1));
        }
    }
    waveAdd(misses, ctr(cnt, part, kCMiss));
}

// The tails of a hand-off bounce, run after k_bounce<..., TAIL> of the same depth and before that depth's shadow
// kernel. LEVEL 1: the traversals its waves handed off, resumed from their records in full waves; with P.tail2 these
// waves hand their own stragglers on in the same way. LEVEL 2: those, resumed to the end. Each is then shaded and
// appended like the bounce's own rays.
template <int TRI, int SL, int WAVES, bool FULL, int LEVEL = 1, bool INST = false>
__global__ void __launch_bounds__(kBlock, WAVES) k_bounce_tail(DevScene S, DevPaths P, const FrameParams* __restrict__ Fp,
                                                               int depth, int smpBase, int K) {
    const FrameParams& F = *Fp;
    __shared__ Group lds[SL * kBlock];
    const int lane = laneId();
    const int gtid = blockIdx.x * kBlock + threadIdx.x;
    LaneStack<Group, SL> st{ (LdsPtr<Group>)(lds + threadIdx.x), (Group*)P.spill + (size_t)gtid * P.spillCap, 0, 0 };
    const int part = (int)(blockIdx.x % kParts), lw = (int)(blockIdx.x / kParts), nw = (int)(gridDim.x / kParts);
    int* cnt = P.cnt + depth * kCntStride;
    const int nt = ctrv(cnt, part, LEVEL == 1 ? kCTail : kCTail2);
    const int inStart = depth == 0 ? partStart(cnt, part) : partStart(cnt - kCntStride, part);
    const int outStart = partStart(cnt, part);
    const size_t first1 = tailFirst(inStart, part, P.tailLanes), first2 = tailFirst(first1, part, P.tailLanes);
    const Vec4* tail = LEVEL == 1 ? P.tail + first1 * (size_t)P.tailStride : P.tail2 + first2 * (size_t)P.tailStride;
    int misses = 0;
    for (int base = lw * kWave; base < nt; base += nw * kWave) {
        const int k = base + lane;
        int i = 0;
        ShadeOut res;
        res.cont = false; res.shadow = false; res.miss = false;
        if (k < nt) {
            Hit h; TravState ts;
            const int slot = resumeRecord(tail + (size_t)k * P.tailStride, h, ts, st);
            Vec3 o, d; float tmin;
            pathRay(P, F, depth, slot, smpBase, K, i, o, d, tmin);
            const RayPre r = resumableRay<TRI, INST>(o, d, tmin);
            if (LEVEL == 1 && P.tail2)
                stepOrHandOff<false, TRI, INST>(S, P, P.tail2, first2, ctr(cnt, part, kCTail2), slot, o, d, tmin, r, h, st, ts);
            else while (!resumableStep<false, TRI, INST>(S, o, d, tmin, r, h, st, ts)) {}
            res = shadeSlot<INST, FULL>(S, P, F, depth, slot, i, smpBase, K, o, d, h);   // nothing if handed on
        }
        misses += appendResults(P, cnt, part, outStart, depth, i, res);
    }
    waveAdd(misses, ctr(cnt, part, kCMiss));
}

// Closest hits of the depth-`depth` queue with persistent lanes and dynamic ray fetching (after Aila & Laine 2009;
// reference: persistentTrace, src/render/wavefront.cu:403-434): a lane whose ray finishes stays idle until at least
// REFILL lanes of its wave are idle, then they all take the next entries of the wave's share of its partition, so
// long rays do not leave the rest of the wave empty. Hits go to H[queue slot] for k_shade.
template <int V, int TRI, int SL, int REFILL>
__global__ void __launch_bounds__(kBlock, 8) k_trace_paths(DevScene S, DevPaths P, const FrameParams* __restrict__ Fp, int depth, int smpBase, int K) {
    const FrameParams& F = *Fp;
    using E = typename StackEntry<V>::type;
    __shared__ E lds[SL > 0 ? SL * kBlock : 1];
    const int gtid = blockIdx.x * kBlock + threadIdx.x;
    LaneStack<E, SL> st{ (LdsPtr<E>)(lds + threadIdx.x), (E*)P.spill + (size_t)gtid * P.spillCap, 0, 0 };
    const int part = (int)(blockIdx.x % kParts), lw = (int)(blockIdx.x / kParts), nw = (int)(gridDim.x / kParts);
    const int* cnt = P.cnt + depth * kCntStride;
    const int n = ctrv(cnt, part, kCQueue);
    const int inStart = depth == 0 ? partStart(cnt, part) : partStart(cnt - kCntStride, part);
    int i = -1;                       // queue slot of this lane's ray, -1 = idle
    RayPre r; Hit h; TravState ts;
    uint32_t f = 0;                   // wave-uniform: entries of the wave's share handed out so far
    bool more = fetchPos(0, lw, nw) < n;
    for (;;) {
        const uint64_t idle = __ballot(i < 0);
        if (more && (int)__builtin_popcountll(idle) >= REFILL) {
            if (i < 0) {
                const uint32_t rank = __builtin_amdgcn_mbcnt_hi((uint32_t)(idle >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)idle, 0u));
                const int k = fetchPos(f + rank, lw, nw);
                if (k < n) {
                    i = inStart + k;
                    int pid; Vec3 o, d; float tmin;
                    pathRay(P, F, depth, i, smpBase, K, pid, o, d, tmin);
                    r = prepareRay<TRI>(o, d, tmin);
                    h.t = 1e30f; h.u = 0.f; h.v = 0.f; h.tri = -1;
                    travBegin(ts, S.root);
                    st.sp = 0;
                }
            }
            f += (uint32_t)__builtin_popcountll(idle);
            more = fetchPos(f, lw, nw) < n;
        }
        if (__ballot(i >= 0) == 0) break;   // no ray in flight: with `more` false the wave's share is done
        if (i >= 0) {
            bool done;
            if constexpr (V == 2) done = stepBVH2<false, TRI>(S.nodes, S.tri, r, h, st, ts);
            else if constexpr (V == 4) done = stepBVH4<false, TRI>(S.nodes, S.tri, r, h, st, ts);
            else done = stepBVH8<false, TRI>(S.nodes, S.tri, r, h, st, ts);
            if (done) { P.H[i] = v4(h.t, h.u, h.v, asFloat(h.tri)); i = -1; }
        }
    }
}


// The fused trace+shade bounce with persistent lanes: a lane whose traversal is over waits (keeping its hit) until
// REFILL lanes of the wave are done (or no lane traces any more), then the done lanes shade together, append their
// results, and take new rays from the partition's queue, while the other lanes keep their traversal state. A wave
// therefore does not idle until its longest ray is done, and shading still runs on many lanes at once. Each path's
// arithmetic is the one of k_bounce (same image); only the order of the appends differs.
template <int V, int TRI, int SL, int WAVES, bool FULL, int REFILL>
__global__ void __launch_bounds__(kBlock, WAVES) k_bounce_refill(DevScene S, DevPaths P, const FrameParams* __restrict__ Fp,
                                                                 int depth, int smpBase, int K) {
    const FrameParams& F = *Fp;
    using E = typename StackEntry<V>::type;
    __shared__ E lds[SL > 0 ? SL * kBlock : 1];
    const int gtid = blockIdx.x * kBlock + threadIdx.x;
    LaneStack<E, SL> st{ (LdsPtr<E>)(lds + threadIdx.x), (E*)P.spill + (size_t)gtid * P.spillCap, 0, 0 };
    const int part = (int)(blockIdx.x % kParts), lw = (int)(blockIdx.x / kParts), nw = (int)(gridDim.x / kParts);
    int* cnt = P.cnt + depth * kCntStride;
    const int n = ctrv(cnt, part, kCQueue);
    const int inStart = depth == 0 ? partStart(cnt, part) : partStart(cnt - kCntStride, part);
    const int outStart = partStart(cnt, part);
    int slot = -1;                    // queue slot of this lane's ray, -1 = idle
    int i = 0;                        // its path id
    bool tracing = false;
    RayPre r; Hit h; TravState ts;
    uint32_t f = 0;                   // wave-uniform: entries of the wave's share handed out so far
    bool more = fetchPos(0, lw, nw) < n;
    int misses = 0;
    for (;;) {
        const uint64_t idle = __ballot(slot < 0);
        if (more && idle) {
            if (slot < 0) {
                const uint32_t rank = __builtin_amdgcn_mbcnt_hi((uint32_t)(idle >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)idle, 0u));
                const int k = fetchPos(f + rank, lw, nw);
                if (k < n) {
                    slot = inStart + k;
                    Vec3 o, d; float tmin;
                    pathRay(P, F, depth, slot, smpBase, K, i, o, d, tmin);
                    r = prepareRay<TRI>(o, d, tmin);
                    h.t = 1e30f; h.u = 0.f; h.v = 0.f; h.tri = -1; h.inst = -1;
                    travBegin(ts, S.root);
                    st.sp = 0;
                    tracing = true;
                }
            }
            f += (uint32_t)__builtin_popcountll(idle);
            more = fetchPos(f, lw, nw) < n;
        }
        if (__ballot(slot >= 0) == 0) break;   // no ray in flight: with `more` false the wave's share is done
        for (;;) {
            if (tracing) {
                bool done;
                if constexpr (V == 2) done = stepBVH2<false, TRI>(S.nodes, S.tri, r, h, st, ts);
                else if constexpr (V == 4) done = stepBVH4<false, TRI>(S.nodes, S.tri, r, h, st, ts);
                else done = stepBVH8<false, TRI>(S.nodes, S.tri, r, h, st, ts);
                if (done) tracing = false;
            }
            const uint64_t tr = __ballot(tracing);
            if (tr == 0 || (more && (int)__builtin_popcountll(__ballot(slot >= 0 && !tracing)) >= REFILL)) break;
        }
        const bool sh = slot >= 0 && !tracing;
        ShadeOut res;
        res.cont = false; res.shadow = false; res.miss = false;
        if (sh) {
            const Vec3 T = pathThroughput(P, depth, slot);
            if constexpr (FULL) {
                const float lastPdf = pathLastPdf(P, depth, slot);
                res = shadeFull<false>(S, P, F, depth, i, i / K, (uint32_t)(smpBase + i % K), r.o, r.d, T, lastPdf, h);
            } else {
                res = shadeLegacy<false>(S, P, F, depth, i, i / K, (uint32_t)(smpBase + i % K), r.o, r.d, T, h);
            }
        }
        misses += appendResults(P, cnt, part, outStart, depth, sh ? i : 0, res);
        if (sh) slot = -1;
    }
    waveAdd(misses, ctr(cnt, part, kCMiss));
}

// Shadow rays of bounce `depth` (partition per XCD as in k_bounce): any hit; unoccluded contributions are added to
// the path's radiance. The any-hit traversal fits 64 registers without spilling: 8 waves per SIMD.
// Two-level scenes keep the TLAS ray and the object-space ray live at once: a budget of 6 waves per SIMD (~84
// registers) is faster than 4 (no spills, less latency hiding) and than 8 (more spills) on the 1,000-instance frame.
constexpr int kInstWaves = 6;
// The full integrator's shading (material blend, light choice, BSDF sample) needs about twice the legacy registers;
// 6 waves per SIMD beat 4 (85 -> 77 ms on a 720p terrain frame) and 8.
constexpr int kFullWaves = 6;

template <int V, int TRI, int SL, bool INST, bool TAIL = false>
__global__ void __launch_bounds__(kBlock, INST ? kInstWaves : 8) k_shadow(DevScene S, DevPaths P, int depth) {
    static_assert(!TAIL || V == 8, "the hand-off resumes BVH8 traversals (two-level scenes: a BVH8 TLAS)");
    using E = typename StackEntry<V>::type;
    __shared__ E lds[SL > 0 ? SL * kBlock : 1];
    const int gtid = blockIdx.x * kBlock + threadIdx.x;
    LaneStack<E, SL> st{ (LdsPtr<E>)(lds + threadIdx.x), (E*)P.spill + (size_t)gtid * P.spillCap, 0, 0 };
    const int part = (int)(blockIdx.x % kParts), lw = (int)(blockIdx.x / kParts), nw = (int)(gridDim.x / kParts);
    int* cnt = P.cnt + depth * kCntStride;
    const int n = ctrv(cnt, part, kCShadow);
    const int start = partStart(cnt, part);
    int unocc = 0;
    for (int base = lw * kWave; base < n; base += nw * kWave) {
        const int k = base + laneId();
        bool vis = false;
        if (k < n) {
            const Vec4 O = P.sO[start + k], D = P.sD[start + k];
            Hit h; h.t = D.w; h.u = 0.f; h.v = 0.f; h.tri = -1; h.inst = -1;
            bool occluded;
            if constexpr (TAIL) {   // hand-off records hold the shadow-queue position as their slot
                traceOrHandOff<true, TRI, INST>(S, P, P.tail, tailFirst(start, part, P.tailLanes), ctr(cnt, part, kCShadowTail),
                                                start + k, xyz(O), xyz(D), 0.f, h, st);
                occluded = h.tri != -1;
            } else {
                occluded = traceRay<V, true, TRI, INST>(S, xyz(O), xyz(D), 0.f, h, st);
            }
            if (!occluded) {
                vis = true;
                const int i = asInt(O.w);
                const Vec4 c = P.sC[start + k];
                Vec4 l = P.L[i]; l.x += c.x; l.y += c.y; l.z += c.z; P.L[i] = l;
            }
        }
        unocc += (int)__builtin_popcountll(__ballot(vis));
    }
    waveAdd(unocc, ctr(cnt, part, kCUnocc));
}

// The tails of a hand-off shadow kernel: its handed-off any-hit traversals resumed in full waves (LEVEL 1, handing
// their own stragglers on to tail2 when it is set; LEVEL 2 resumes those to the end).
template <int TRI, int SL, int LEVEL = 1, bool INST = false>
__global__ void __launch_bounds__(kBlock, INST ? kInstWaves : 8) k_shadow_tail(DevScene S, DevPaths P, int depth) {
    __shared__ Group lds[SL * kBlock];
    const int gtid = blockIdx.x * kBlock + threadIdx.x;
    LaneStack<Group, SL> st{ (LdsPtr<Group>)(lds + threadIdx.x), (Group*)P.spill + (size_t)gtid * P.spillCap, 0, 0 };
    const int part = (int)(blockIdx.x % kParts), lw = (int)(blockIdx.x / kParts), nw = (int)(gridDim.x / kParts);
    int* cnt = P.cnt + depth * kCntStride;
    const int nt = ctrv(cnt, part, LEVEL == 1 ? kCShadowTail : kCShadowTail2);
    const int start = partStart(cnt, part);
    const size_t first1 = tailFirst(start, part, P.tailLanes), first2 = tailFirst(first1, part, P.tailLanes);
    const Vec4* tail = LEVEL == 1 ? P.tail + first1 * (size_t)P.tailStride : P.tail2 + first2 * (size_t)P.tailStride;
    int unocc = 0;
    for (int base = lw * kWave; base < nt; base += nw * kWave) {
        const int k = base + laneId();
        bool vis = false;
        if (k < nt) {
            Hit h; TravState ts;
            const int pos = resumeRecord(tail + (size_t)k * P.tailStride, h, ts, st);
            const Vec4 O = P.sO[pos], D = P.sD[pos];
            const Vec3 so = xyz(O), sd = xyz(D);
            const RayPre r = resumableRay<TRI, INST>(so, sd, 0.f);
            if (LEVEL == 1 && P.tail2)
                stepOrHandOff<true, TRI, INST>(S, P, P.tail2, first2, ctr(cnt, part, kCShadowTail2), pos, so, sd, 0.f, r, h, st, ts);
            else while (!resumableStep<true, TRI, INST>(S, so, sd, 0.f, r, h, st, ts)) {}
            if (h.tri == -1) {
                vis = true;
                const int i = asInt(O.w);
                const Vec4 c = P.sC[pos];
                Vec4 l = P.L[i]; l.x += c.x; l.y += c.y; l.z += c.z; P.L[i] = l;
            }
        }
        unocc += (int)__builtin_popcountll(__ballot(vis));
    }
    waveAdd(unocc, ctr(cnt, part, kCUnocc));
}

// Shadow rays with persistent lanes and dynamic ray fetching (as k_trace_paths).
template <int V, int TRI, int SL, int REFILL>
__global__ void __launch_bounds__(kBlock, 8) k_shadow_refill(DevScene S, DevPaths P, int depth) {
    using E = typename StackEntry<V>::type;
    __shared__ E lds[SL > 0 ? SL * kBlock : 1];
    const int gtid = blockIdx.x * kBlock + threadIdx.x;
    LaneStack<E, SL> st{ (LdsPtr<E>)(lds + threadIdx.x), (E*)P.spill + (size_t)gtid * P.spillCap, 0, 0 };
    const int part = (int)(blockIdx.x % kParts), lw = (int)(blockIdx.x / kParts), nw = (int)(gridDim.x / kParts);
    int* cnt = P.cnt + depth * kCntStride;
    const int n = ctrv(cnt, part, kCShadow);
    const int start = partStart(cnt, part);
    int k = -1;                       // shadow-queue entry of this lane's ray, -1 = idle
    RayPre r; Hit h; TravState ts;
    uint32_t f = 0;
    bool more = fetchPos(0, lw, nw) < n;
    int unocc = 0;
    for (;;) {
        const uint64_t idle = __ballot(k < 0);
        if (more && (int)__builtin_popcountll(idle) >= REFILL) {
            if (k < 0) {
                const uint32_t rank = __builtin_amdgcn_mbcnt_hi((uint32_t)(idle >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)idle, 0u));
                const int kk = fetchPos(f + rank, lw, nw);
                if (kk < n) {
                    k = kk;
                    const Vec4 O = P.sO[start + k], D = P.sD[start + k];
                    r = prepareRay<TRI>(xyz(O), xyz(D), 0.f);
                    h.t = D.w; h.u = 0.f; h.v = 0.f; h.tri = -1;
                    travBegin(ts, S.root);
                    st.sp = 0;
                }
            }
            f += (uint32_t)__builtin_popcountll(idle);
            more = fetchPos(f, lw, nw) < n;
        }
        if (__ballot(k >= 0) == 0) break;
        if (k >= 0) {
            bool done;
            if constexpr (V == 2) done = stepBVH2<true, TRI>(S.nodes, S.tri, r, h, st, ts);
            else if constexpr (V == 4) done = stepBVH4<true, TRI>(S.nodes, S.tri, r, h, st, ts);
            else done = stepBVH8<true, TRI>(S.nodes, S.tri, r, h, st, ts);
            if (done) {
                if (h.tri < 0) {
                    unocc++;
                    const int i = asInt(P.sO[start + k].w);
                    const Vec4 c = P.sC[start + k];
                    Vec4 l = P.L[i]; l.x += c.x; l.y += c.y; l.z += c.z; P.L[i] = l;
                }
                k = -1;
            }
        }
    }
    // one atomic per wave for the unoccluded count
    int sum = unocc;
    for (int off = 32; off > 0; off >>= 1) sum += __shfl_xor(sum, off);
    waveAdd(sum, ctr(cnt, part, kCUnocc));
}

// Closest or any hit of explicit rays (parity and kernel-level entry point); two-level scenes also report the hit
// instance (-1 on a miss) when hitInst is given.
template <int V, int TRI, bool ANY, int SL, bool INST>
__global__ void __launch_bounds__(kBlock, INST ? kInstWaves : 8) k_trace(DevScene S, const Vec4* __restrict__ rays, int n, Vec4* __restrict__ hits,
                                                     int* __restrict__ hitInst, void* spill, int spillCap) {
    using E = typename StackEntry<V>::type;
    __shared__ E lds[SL > 0 ? SL * kBlock : 1];
    const int gtid = blockIdx.x * kBlock + threadIdx.x;
    LaneStack<E, SL> st{ (LdsPtr<E>)(lds + threadIdx.x), (E*)spill + (size_t)gtid * spillCap, 0, 0 };
    for (int k = gtid; k - (int)threadIdx.x < n; k += gridDim.x * kBlock) {
        if (k >= n) continue;
        const Vec4 O = rays[2 * (size_t)k], D = rays[2 * (size_t)k + 1];
        Hit h; h.t = D.w; h.u = 0.f; h.v = 0.f; h.tri = -1; h.inst = -1;
        traceRay<V, ANY, TRI, INST>(S, xyz(O), xyz(D), O.w, h, st);
        hits[k] = v4(h.t, h.u, h.v, asFloat(h.tri));
        if (hitInst) hitInst[k] = h.tri >= 0 ? h.inst : -1;
    }
}

// ------------------------------------------------------------------------------------------------ batched jobs
// A render job of a batched pass: its scene, frame parameters and image, and its slices of the pass's concatenated
// path ids (path pathBase + pixel * K + sample) and 64-pixel reduction groups.
struct DevJob {
    DevScene S;
    FrameParams F;
    Vec4* accum;        // per-pixel radiance sums of the job (N float4)
    int pathBase;       // first path id of the job in this pass
    int N, K, smpBase;  // pixels, samples in flight, sample index of the pass's first sample
    int grpBase;        // first 64-pixel reduction group of the job
};
constexpr int kMaxJobs = 256;

// The job table of a pass: a header the kernels read on the device (so that a pass is a fixed launch sequence that
// can be captured once in a graph and replayed with any jobs), then the jobs.
struct DevJobTable {
    int nJobs;          // jobs in the pass
    int groups;         // 64-pixel reduction groups over all jobs
    long long paths;    // paths over all jobs
    int pad[12];        // 64-byte header
    DevJob jobs[kMaxJobs];
};

// the job of path id i: the last job whose first id is <= i (bases ascending, in LDS)
__device__ __forceinline__ int jobOf(const int* base, int n, int i) {
    int lo = 0, hi = n;
    while (hi - lo > 1) { const int mid = (lo + hi) >> 1; if (base[mid] <= i) lo = mid; else hi = mid; }
    return lo;
}

// The ray of entry `slot` of a pass's depth-`depth` queue (depth 0: the camera ray of path id slot): its path id, its
// job (the last whose first id is <= it), pixel and sample index.
struct JobEntry { int i, j, pix; uint32_t smp; };
__device__ __forceinline__ JobEntry jobEntry(const DevJob* __restrict__ jobs, const int* base, int nJobs, const DevPaths& P,
                                             int depth, int slot, Vec3& o, Vec3& d, float& tmin) {
    JobEntry e;
    Vec4 Or = v4(0.f, 0.f, 0.f, 0.f);
    if (depth == 0) e.i = slot;
    else { Or = P.O[depth & 1][slot]; e.i = asInt(Or.w); }
    e.j = jobOf(base, nJobs, e.i);
    const DevJob& J = jobs[e.j];
    const int li = e.i - base[e.j], Kj = J.K;
    e.pix = li / Kj;
    e.smp = (uint32_t)(J.smpBase + li % Kj);
    if (depth == 0) { cameraRay(J.F, e.pix, e.smp, o, d); tmin = J.F.nearT; }
    else { o = xyz(Or); d = xyz(P.D[depth & 1][slot]); tmin = 0.f; }
    return e;
}
// shading of a pass entry's traced ray with its job's scene and parameters (nothing for a handed-off traversal)
template <bool FULL>
__device__ __forceinline__ ShadeOut jobShade(const DevJob& J, const DevPaths& P, int depth, int slot, const JobEntry& e,
                                             Vec3 o, Vec3 d, const Hit& h) {
    const Vec3 T = pathThroughput(P, depth, slot);
    if constexpr (FULL) {
        const float lastPdf = pathLastPdf(P, depth, slot);
        return shadeFull<false>(J.S, P, J.F, depth, e.i, e.pix, e.smp, o, d, T, lastPdf, h);
    } else {
        return shadeLegacy<false>(J.S, P, J.F, depth, e.i, e.pix, e.smp, o, d, T, h);
    }
}

// The job table's path bases into LDS (for jobOf); every thread of the block meets the barrier.
__device__ __forceinline__ void loadJobBases(const DevJob* __restrict__ jobs, int nJobs, int* base) {
    for (int j = threadIdx.x; j < nJobs; j += kBlock) base[j] = jobs[j].pathBase;
    __syncthreads();
}

// k_bounce over the paths of many jobs at once (one launch per bounce for the whole batch): each path finds its job,
// and with it the scene it traverses and the parameters it is shaded with. A wave may hold paths of several jobs.
// TAIL (BVH8): the straggler hand-off of k_bounce: once fewer than P.tailLanes lanes of a wave are still tracing, they
// write their traversals as tail records (k_bounce_jobs_tail resumes them with each record's own job and scene).
template <int V, int TRI, int SL, int WAVES, bool FULL = false, bool TAIL = false>
__global__ void __launch_bounds__(kBlock, WAVES) k_bounce_jobs(const DevJobTable* __restrict__ T, DevPaths P, int depth) {
    static_assert(!TAIL || V == 8, "the hand-off resumes BVH8 traversals");
    using E = typename StackEntry<V>::type;
    __shared__ E lds[SL > 0 ? SL * kBlock : 1];
    __shared__ int base[kMaxJobs];
    const int nJobs = T->nJobs;
    const DevJob* __restrict__ jobs = T->jobs;
    loadJobBases(jobs, nJobs, base);
    const int lane = laneId();
    const int gtid = blockIdx.x * kBlock + threadIdx.x;
    LaneStack<E, SL> st{ (LdsPtr<E>)(lds + threadIdx.x), (E*)P.spill + (size_t)gtid * P.spillCap, 0, 0 };
    const int part = (int)(blockIdx.x % kParts), lw = (int)(blockIdx.x / kParts), nw = (int)(gridDim.x / kParts);
    int* cnt = P.cnt + depth * kCntStride;
    const int n = ctrv(cnt, part, kCQueue);
    const int inStart = depth == 0 ? partStart(cnt, part) : partStart(cnt - kCntStride, part);
    const int outStart = partStart(cnt, part);
    int misses = 0;
    for (int b0 = lw * kWave; b0 < n; b0 += nw * kWave) {
        const int k = b0 + lane;
        int i = 0;
        ShadeOut res;
        res.cont = false; res.shadow = false; res.miss = false;
        if (k < n) {
            const int slot = inStart + k;
            Vec3 o, d; float tmin;
            const JobEntry e = jobEntry(jobs, base, nJobs, P, depth, slot, o, d, tmin);
            const DevJob& J = jobs[e.j];
            i = e.i;
            Hit h; h.t = 1e30f; h.u = 0.f; h.v = 0.f; h.tri = -1; h.inst = -1;
            if constexpr (TAIL)
                traceOrHandOff<false, TRI, false>(J.S, P, P.tail, tailFirst(inStart, part, P.tailLanes), ctr(cnt, part, kCTail),
                                                  slot, o, d, tmin, h, st);
            else
                traceRay<V, false, TRI, false>(J.S, o, d, tmin, h, st);
            res = jobShade<FULL>(J, P, depth, slot, e, o, d, h);
        }
        misses += appendResults(P, cnt, part, outStart, depth, i, res);
    }
    waveAdd(misses, ctr(cnt, part, kCMiss));
}

// The tail of a hand-off pass bounce (run after k_bounce_jobs<..., TAIL> of the same depth, before its shadow
// kernel): the handed-off traversals resumed from their records in full waves, each with its job's scene, to the end;
// then shaded and appended like the bounce's own rays. One level: a second one (as the rolling steps') made the
// one-job-at-a-time stream 2.7 % slower.
template <int TRI, int SL, int WAVES, bool FULL>
__global__ void __launch_bounds__(kBlock, WAVES) k_bounce_jobs_tail(const DevJobTable* __restrict__ T, DevPaths P, int depth) {
    __shared__ Group lds[SL * kBlock];
    __shared__ int base[kMaxJobs];
    const int nJobs = T->nJobs;
    const DevJob* __restrict__ jobs = T->jobs;
    loadJobBases(jobs, nJobs, base);
    const int lane = laneId();
    const int gtid = blockIdx.x * kBlock + threadIdx.x;
    LaneStack<Group, SL> st{ (LdsPtr<Group>)(lds + threadIdx.x), (Group*)P.spill + (size_t)gtid * P.spillCap, 0, 0 };
    const int part = (int)(blockIdx.x % kParts), lw = (int)(blockIdx.x / kParts), nw = (int)(gridDim.x / kParts);
    int* cnt = P.cnt + depth * kCntStride;
    const int nt = ctrv(cnt, part, kCTail);
    const int inStart = depth == 0 ? partStart(cnt, part) : partStart(cnt - kCntStride, part);
    const int outStart = partStart(cnt, part);
    const Vec4* tail = P.tail + tailFirst(inStart, part, P.tailLanes) * (size_t)P.tailStride;
    int misses = 0;
    for (int b0 = lw * kWave; b0 < nt; b0 += nw * kWave) {
        const int k = b0 + lane;
        int i = 0;
        ShadeOut res;
        res.cont = false; res.shadow = false; res.miss = false;
        if (k < nt) {
            Hit h; TravState ts;
            const int slot = resumeRecord(tail + (size_t)k * P.tailStride, h, ts, st);
            Vec3 o, d; float tmin;
            const JobEntry e = jobEntry(jobs, base, nJobs, P, depth, slot, o, d, tmin);
            const DevJob& J = jobs[e.j];
            i = e.i;
            const RayPre r = resumableRay<TRI, false>(o, d, tmin);
            while (!resumableStep<false, TRI, false>(J.S, o, d, tmin, r, h, st, ts)) {}
            res = jobShade<FULL>(J, P, depth, slot, e, o, d, h);
        }
        misses += appendResults(P, cnt, part, outStart, depth, i, res);
    }
    waveAdd(misses, ctr(cnt, part, kCMiss));
}

// k_shadow over the shadow rays of many jobs: the ray's path id names its job and scene (TAIL: with the hand-off)
template <int V, int TRI, int SL, bool TAIL = false>
__global__ void __launch_bounds__(kBlock, 8) k_shadow_jobs(const DevJobTable* __restrict__ T, DevPaths P, int depth) {
    static_assert(!TAIL || V == 8, "the hand-off resumes BVH8 traversals");
    using E = typename StackEntry<V>::type;
    __shared__ E lds[SL > 0 ? SL * kBlock : 1];
    __shared__ int base[kMaxJobs];
    const int nJobs = T->nJobs;
    const DevJob* __restrict__ jobs = T->jobs;
    loadJobBases(jobs, nJobs, base);
    const int gtid = blockIdx.x * kBlock + threadIdx.x;
    LaneStack<E, SL> st{ (LdsPtr<E>)(lds + threadIdx.x), (E*)P.spill + (size_t)gtid * P.spillCap, 0, 0 };
    const int part = (int)(blockIdx.x % kParts), lw = (int)(blockIdx.x / kParts), nw = (int)(gridDim.x / kParts);
    int* cnt = P.cnt + depth * kCntStride;
    const int n = ctrv(cnt, part, kCShadow);
    const int start = partStart(cnt, part);
    int unocc = 0;
    for (int b0 = lw * kWave; b0 < n; b0 += nw * kWave) {
        const int k = b0 + laneId();
        bool vis = false;
        if (k < n) {
            const Vec4 O = P.sO[start + k], D = P.sD[start + k];
            const int i = asInt(O.w);
            const DevJob& J = jobs[jobOf(base, nJobs, i)];
            Hit h; h.t = D.w; h.u = 0.f; h.v = 0.f; h.tri = -1; h.inst = -1;
            bool occluded;
            if constexpr (TAIL) {   // records hold the shadow-queue position as their slot
                traceOrHandOff<true, TRI, false>(J.S, P, P.tail, tailFirst(start, part, P.tailLanes), ctr(cnt, part, kCShadowTail),
                                                 start + k, xyz(O), xyz(D), 0.f, h, st);
                occluded = h.tri != -1;
            } else {
                occluded = traceRay<V, true, TRI, false>(J.S, xyz(O), xyz(D), 0.f, h, st);
            }
            if (!occluded) {
                vis = true;
                const Vec4 c = P.sC[start + k];
                Vec4 l = P.L[i]; l.x += c.x; l.y += c.y; l.z += c.z; P.L[i] = l;
            }
        }
        unocc += (int)__builtin_popcountll(__ballot(vis));
    }
    waveAdd(unocc, ctr(cnt, part, kCUnocc));
}

// The tail of a hand-off batch shadow kernel: its handed-off any-hit traversals resumed to the end, each with the
// scene of its path's job. Pass (ROLL false, x = depth): counter block `depth`, shadow rays from the prefix of the
// depth's queue; rolling step (ROLL true, x = step): block step & 1, from the prefix of the step's inputs.
template <int TRI, int SL, bool ROLL, class Table>
__global__ void __launch_bounds__(kBlock, 8) k_shadow_jobs_tail(const Table* __restrict__ T, DevPaths P, int x) {
    __shared__ Group lds[SL * kBlock];
    __shared__ int base[kMaxJobs];
    const int nJobs = T->nJobs;
    const DevJob* __restrict__ jobs = T->jobs;
    loadJobBases(jobs, nJobs, base);
    const int gtid = blockIdx.x * kBlock + threadIdx.x;
    LaneStack<Group, SL> st{ (LdsPtr<Group>)(lds + threadIdx.x), (Group*)P.spill + (size_t)gtid * P.spillCap, 0, 0 };
    const int part = (int)(blockIdx.x % kParts), lw = (int)(blockIdx.x / kParts), nw = (int)(gridDim.x / kParts);
    int* cnt = P.cnt + (ROLL ? (x & 1) : x) * kCntStride;
    const int nt = ctrv(cnt, part, kCShadowTail);
    const int start = partStart(cnt, part, ROLL ? kCIn : kCQueue);
    const Vec4* tail = P.tail + tailFirst(start, part, P.tailLanes) * (size_t)P.tailStride;
    int unocc = 0;
    for (int b0 = lw * kWave; b0 < nt; b0 += nw * kWave) {
        const int k = b0 + laneId();
        bool vis = false;
        if (k < nt) {
            Hit h; TravState ts;
            const int pos = resumeRecord(tail + (size_t)k * P.tailStride, h, ts, st);
            const Vec4 O = P.sO[pos], D = P.sD[pos];
            const int i = asInt(O.w);
            const DevJob& J = jobs[jobOf(base, nJobs, i)];
            const Vec3 so = xyz(O), sd = xyz(D);
            const RayPre r = resumableRay<TRI, false>(so, sd, 0.f);
            while (!resumableStep<true, TRI, false>(J.S, so, sd, 0.f, r, h, st, ts)) {}
            if (h.tri == -1) {
                vis = true;
                const Vec4 c = P.sC[pos];
                Vec4 l = P.L[i]; l.x += c.x; l.y += c.y; l.z += c.z; P.L[i] = l;
            }
        }
        unocc += (int)__builtin_popcountll(__ballot(vis));
    }
    waveAdd(unocc, ctr(cnt, part, kCUnocc));
}

// ------------------------------------------------------------------------------------------------ rolling batches
// A rolling batch runs STEPS rather than passes: every step is one trace+shade launch and one shadow launch over every
// live path of every live job, whatever its depth, plus the camera rays of the jobs admitted at that step. A path
// carries its depth (T.w), so its arithmetic, random numbers and radiance are those of the job rendered alone; a job
// admitted at step s has traced all its rays at step s + bounces + 1 and is reduced right after that step. New jobs
// therefore join the very next launch instead of waiting for a pass to end, and the deep bounces of old jobs share
// launches with the camera rays of new ones (no per-pass tail of nearly empty launches).
//
// Step s reads its continuations from ray set s & 1 (appended by step s - 1) and appends to set (s + 1) & 1; its
// counters are block s & 1 of a ring of two. Partition x's input is the previous step's continuations of partition x,
// then its share of the step's camera rays; it appends at the start given by the prefix of the partitions' inputs.
struct DevStepTable {
    int nJobs;            // live jobs, path bases ascending
    int nNew;             // jobs whose camera rays this step traces
    int nDone;            // jobs reduced after this step
    int doneGroups;       // 64-pixel reduction groups over the done jobs
    long long camPaths;   // camera rays of this step
    int pad[10];          // 64-byte header
    int newJob[kMaxJobs];    // new jobs (indices into jobs), in camera-index order
    int camFirst[kMaxJobs];  // first camera index of each new job (ascending)
    int doneJob[kMaxJobs];   // done jobs (indices into jobs)
    int doneGrp[kMaxJobs];   // first reduction group of each done job (ascending)
    DevJob jobs[kMaxJobs];
};

// the new job (index into newJob / camFirst) of camera index c: the last one whose first index is <= c
__device__ __forceinline__ int camJobOf(const DevStepTable* __restrict__ T, int c) {
    int lo = 0, hi = T->nNew;
    while (hi - lo > 1) { const int mid = (lo + hi) >> 1; if (T->camFirst[mid] <= c) lo = mid; else hi = mid; }
    return lo;
}

// The ray of entry `code` of a rolling step: the continuation at slot `code` of ray set inSet (code >= 0) or the
// camera ray of the step's camera index -(code + 1); its path id, job, pixel and sample index.
struct RollEntry { int i, j, pix; uint32_t smp; };
__device__ __forceinline__ RollEntry rollEntry(const DevStepTable* __restrict__ T, const int* base, int nJobs, const DevPaths& P,
                                               int inSet, int code, Vec3& o, Vec3& d, float& tmin) {
    RollEntry e;
    Vec4 Or = v4(0.f, 0.f, 0.f, 0.f);
    if (code >= 0) {
        Or = P.O[inSet][code];
        e.i = asInt(Or.w);
        e.j = jobOf(base, nJobs, e.i);
    } else {
        const int c = -(code + 1);
        const int q = camJobOf(T, c);
        e.j = T->newJob[q];
        e.i = base[e.j] + (c - T->camFirst[q]);
    }
    const DevJob& J = T->jobs[e.j];
    const int li = e.i - base[e.j], Kj = J.K;
    e.pix = li / Kj;
    e.smp = (uint32_t)(J.smpBase + li % Kj);
    if (code < 0) { cameraRay(J.F, e.pix, e.smp, o, d); tmin = J.F.nearT; }
    else { o = xyz(Or); d = xyz(P.D[inSet][code]); tmin = 0.f; }
    return e;
}
// shading of a rolling entry's traced ray; `depth` is the path's (0 for camera rays). Nothing for a handed-off one.
template <bool FULL>
__device__ __forceinline__ ShadeOut rollShade(const DevJob& J, const DevPaths& P, int inSet, int code, const RollEntry& e,
                                              Vec3 o, Vec3 d, const Hit& h, int& depth) {
    Vec3 Tp = v3(1.f, 1.f, 1.f);
    float lastPdf = 0.f;
    depth = 0;
    if (code >= 0) { const Vec4 Tv = P.T[inSet][code]; Tp = xyz(Tv); depth = asInt(Tv.w); lastPdf = P.D[inSet][code].w; }
    if constexpr (FULL) return shadeFull<false>(J.S, P, J.F, depth, e.i, e.pix, e.smp, o, d, Tp, lastPdf, h);
    else return shadeLegacy<false>(J.S, P, J.F, depth, e.i, e.pix, e.smp, o, d, Tp, h);
}

// TAIL (BVH8): the straggler hand-off of k_bounce; a record's slot is the entry's code (camera rays negative), and
// k_bounce_roll_tail resumes the records of the step.
template <int V, int TRI, int SL, int WAVES, bool FULL = false, bool LANE = false, bool TAIL = false>
__global__ void __launch_bounds__(kBlock, WAVES) k_bounce_roll(const DevStepTable* __restrict__ T, DevPaths P, int step) {
    static_assert(!TAIL || V == 8, "the hand-off resumes BVH8 traversals");
    using E = typename StackEntry<V>::type;
    __shared__ E lds[SL > 0 ? SL * kBlock : 1];
    __shared__ int base[kMaxJobs];
    const int nJobs = T->nJobs;
    const DevJob* __restrict__ jobs = T->jobs;
    loadJobBases(jobs, nJobs, base);
    const int lane = laneId();
    const int gtid = blockIdx.x * kBlock + threadIdx.x;
    LaneStack<E, SL> st{ (LdsPtr<E>)(lds + threadIdx.x), (E*)P.spill + (size_t)gtid * P.spillCap, 0, 0 };
    const int part = (int)(blockIdx.x % kParts), lw = (int)(blockIdx.x / kParts), nw = (int)(gridDim.x / kParts);
    int* blk = P.cnt + (step & 1) * kCntStride;
    const int* prev = P.cnt + ((step + 1) & 1) * kCntStride;
    const int n = ctrv(blk, part, kCIn);
    const int nCont = ctrv(prev, part, kCQueue);
    const int contStart = partStart(prev, part, kCIn);
    const int outStart = partStart(blk, part, kCIn);
    const int camLo = (int)((long long)part * T->camPaths / kParts);
    const int inSet = step & 1, outSet = (step + 1) & 1;
    int misses = 0, camMisses = 0;
    for (int b0 = lw * kWave; b0 < n; b0 += nw * kWave) {
        const int k = b0 + lane;
        int i = 0, depth = 0;
        bool cam = false;
        ShadeOut res;
        res.cont = false; res.shadow = false; res.miss = false;
        if (k < n) {
            const int code = k < nCont ? contStart + k : -(camLo + (k - nCont)) - 1;
            cam = code < 0;
            Vec3 o, d; float tmin;
            const RollEntry e = rollEntry(T, base, nJobs, P, inSet, code, o, d, tmin);
            const DevJob& J = jobs[e.j];
            i = e.i;
            Hit h; h.t = 1e30f; h.u = 0.f; h.v = 0.f; h.tri = -1; h.inst = -1;
            if constexpr (TAIL)
                traceOrHandOff<false, TRI, false>(J.S, P, P.tail, tailFirst(outStart, part, P.tailLanes), ctr(blk, part, kCTail),
                                                  code, o, d, tmin, h, st);
            else
                traceRay<V, false, TRI, false>(J.S, o, d, tmin, h, st);
            res = rollShade<FULL>(J, P, inSet, code, e, o, d, h, depth);
        }
        camMisses += (int)__builtin_popcountll(__ballot(cam && res.miss));
        misses += appendTo<LANE>(P, ctr(blk, part, kCQueue), ctr(blk, part, kCShadow), outStart, outSet, depth + 1, i, res);
    }
    waveAdd(misses, ctr(blk, part, kCMiss));
    waveAdd(camMisses, ctr(blk, part, kCMissCam));
}

// The tails of a hand-off rolling step (after k_bounce_roll<..., TAIL>, before the step's shadow kernel), as those of a
// pass: LEVEL 1 resumes the step's handed-off traversals in full waves, each with its job's scene, handing its own
// stragglers on when P.tail2 is set; LEVEL 2 resumes those to the end. Each is shaded and appended like the step's own.
template <int TRI, int SL, int WAVES, bool FULL, bool LANE, int LEVEL = 1>
__global__ void __launch_bounds__(kBlock, WAVES) k_bounce_roll_tail(const DevStepTable* __restrict__ T, DevPaths P, int step) {
    __shared__ Group lds[SL * kBlock];
    __shared__ int base[kMaxJobs];
    const int nJobs = T->nJobs;
    const DevJob* __restrict__ jobs = T->jobs;
    loadJobBases(jobs, nJobs, base);
    const int lane = laneId();
    const int gtid = blockIdx.x * kBlock + threadIdx.x;
    LaneStack<Group, SL> st{ (LdsPtr<Group>)(lds + threadIdx.x), (Group*)P.spill + (size_t)gtid * P.spillCap, 0, 0 };
    const int part = (int)(blockIdx.x % kParts), lw = (int)(blockIdx.x / kParts), nw = (int)(gridDim.x / kParts);
    int* blk = P.cnt + (step & 1) * kCntStride;
    const int nt = ctrv(blk, part, LEVEL == 1 ? kCTail : kCTail2);
    const int outStart = partStart(blk, part, kCIn);
    const int inSet = step & 1, outSet = (step + 1) & 1;
    const size_t first1 = tailFirst(outStart, part, P.tailLanes), first2 = tailFirst(first1, part, P.tailLanes);
    const Vec4* tail = LEVEL == 1 ? P.tail + first1 * (size_t)P.tailStride : P.tail2 + first2 * (size_t)P.tailStride;
    int misses = 0, camMisses = 0;
    for (int b0 = lw * kWave; b0 < nt; b0 += nw * kWave) {
        const int k = b0 + lane;
        int i = 0, depth = 0;
        bool cam = false;
        ShadeOut res;
        res.cont = false; res.shadow = false; res.miss = false;
        if (k < nt) {
            Hit h; TravState ts;
            const int code = resumeRecord(tail + (size_t)k * P.tailStride, h, ts, st);
            cam = code < 0;
            Vec3 o, d; float tmin;
            const RollEntry e = rollEntry(T, base, nJobs, P, inSet, code, o, d, tmin);
            const DevJob& J = jobs[e.j];
            i = e.i;
            const RayPre r = resumableRay<TRI, false>(o, d, tmin);
            if (LEVEL == 1 && P.tail2)
                stepOrHandOff<false, TRI, false>(J.S, P, P.tail2, first2, ctr(blk, part, kCTail2), code, o, d, tmin, r, h, st, ts);
            else while (!resumableStep<false, TRI, false>(J.S, o, d, tmin, r, h, st, ts)) {}
            res = rollShade<FULL>(J, P, inSet, code, e, o, d, h, depth);   // nothing if handed on
        }
        camMisses += (int)__builtin_popcountll(__ballot(cam && res.miss));
        misses += appendTo<LANE>(P, ctr(blk, part, kCQueue), ctr(blk, part, kCShadow), outStart, outSet, depth + 1, i, res);
    }
    waveAdd(misses, ctr(blk, part, kCMiss));
    waveAdd(camMisses, ctr(blk, part, kCMissCam));
}

// the step's shadow rays (as k_shadow_jobs, counters of block step & 1; TAIL: with the hand-off)
template <int V, int TRI, int SL, bool TAIL = false>
__global__ void __launch_bounds__(kBlock, 8) k_shadow_roll(const DevStepTable* __restrict__ T, DevPaths P, int step) {
    static_assert(!TAIL || V == 8, "the hand-off resumes BVH8 traversals");
    using E = typename StackEntry<V>::type;
    __shared__ E lds[SL > 0 ? SL * kBlock : 1];
    __shared__ int base[kMaxJobs];
    const int nJobs = T->nJobs;
    const DevJob* __restrict__ jobs = T->jobs;
    loadJobBases(jobs, nJobs, base);
    const int gtid = blockIdx.x * kBlock + threadIdx.x;
    LaneStack<E, SL> st{ (LdsPtr<E>)(lds + threadIdx.x), (E*)P.spill + (size_t)gtid * P.spillCap, 0, 0 };
    const int part = (int)(blockIdx.x % kParts), lw = (int)(blockIdx.x / kParts), nw = (int)(gridDim.x / kParts);
    int* blk = P.cnt + (step & 1) * kCntStride;
    const int n = ctrv(blk, part, kCShadow);
    const int start = partStart(blk, part, kCIn);
    int unocc = 0;
    for (int b0 = lw * kWave; b0 < n; b0 += nw * kWave) {
        const int k = b0 + laneId();
        bool vis = false;
        if (k < n) {
            const Vec4 O = P.sO[start + k], D = P.sD[start + k];
            const int i = asInt(O.w);
            const DevJob& J = jobs[jobOf(base, nJobs, i)];
            Hit h; h.t = D.w; h.u = 0.f; h.v = 0.f; h.tri = -1; h.inst = -1;
            bool occluded;
            if constexpr (TAIL) {
                traceOrHandOff<true, TRI, false>(J.S, P, P.tail, tailFirst(start, part, P.tailLanes), ctr(blk, part, kCShadowTail),
                                                 start + k, xyz(O), xyz(D), 0.f, h, st);
                occluded = h.tri != -1;
            } else {
                occluded = traceRay<V, true, TRI, false>(J.S, xyz(O), xyz(D), 0.f, h, st);
            }
            if (!occluded) {
                vis = true;
                const Vec4 c = P.sC[start + k];
                Vec4 l = P.L[i]; l.x += c.x; l.y += c.y; l.z += c.z; P.L[i] = l;
            }
        }
        unocc += (int)__builtin_popcountll(__ballot(vis));
    }
    waveAdd(unocc, ctr(blk, part, kCUnocc));
}

// ------------------------------------------------------------------------------------------------------- dispatch
// The frame parameters (camera, lights, filter table: ~400 bytes) are read through a pointer: as a by-value argument
// they occupied scalar registers for the whole kernel and made the trace+shade kernel spill (68 -> 24 bytes of
// scratch per lane, +3.6 % frame throughput on the 720p terrain frame).
using BounceFn = void (*)(DevScene, DevPaths, const FrameParams*, int, int, int);
using FirstFn = void (*)(DevScene, DevPaths, const FrameParams*, int, int);
using ShadowFn = void (*)(DevScene, DevPaths, int);
using TraceFn = void (*)(DevScene, const Vec4*, int, Vec4*, int*, void*, int);
using JobsFn = void (*)(const DevJobTable*, DevPaths, int);
using RollFn = void (*)(const DevStepTable*, DevPaths, int);

// the kernels of one (layout, triangle test, LDS stack) configuration
struct Kernels {
    BounceFn bounce = nullptr;                          // fused trace + shade
    BounceFn tracePaths = nullptr, tracePlain = nullptr; // split: closest-hit trace with / without refill (then k_shade)
    ShadowFn shadow = nullptr, shadowRefill = nullptr;
    TraceFn traceC = nullptr, traceA = nullptr;
    // two-level scenes (fused bounce kernel only)
    BounceFn bounceInst = nullptr;
    ShadowFn shadowInst = nullptr;
    TraceFn traceCInst = nullptr, traceAInst = nullptr;
    // batched jobs
    JobsFn bounceJobs = nullptr, shadowJobs = nullptr;
    // full material integrator (fused bounce kernel; shadow rays use the legacy shadow kernels)
    BounceFn bounceFull = nullptr, bounceFullInst = nullptr;
    // fused trace + shade with persistent lanes (legacy and full integrator)
    BounceFn bounceRefill = nullptr, bounceRefillFull = nullptr;
    JobsFn bounceJobsFull = nullptr;
    // rolling batches
    RollFn bounceRoll = nullptr, bounceRollFull = nullptr, shadowRoll = nullptr;
    RollFn bounceRollLane = nullptr, bounceRollFullLane = nullptr;   // per-lane compaction
    // straggler hand-off (BVH8, single-level scenes): the bounce kernels that hand off and their tail kernels
    BounceFn bounceHand = nullptr, bounceTail = nullptr, bounceHandFull = nullptr, bounceTailFull = nullptr;
    ShadowFn shadowHand = nullptr, shadowTail = nullptr;
    BounceFn bounceTail2 = nullptr, bounceTail2Full = nullptr;   // second hand-off level
    ShadowFn shadowTail2 = nullptr;
    // the same for two-level scenes with a BVH8 TLAS (hand-off between TLAS node steps)
    BounceFn bounceHandInst = nullptr, bounceHandFullInst = nullptr, bounceTailInst = nullptr, bounceTailFullInst = nullptr;
    BounceFn bounceTail2Inst = nullptr, bounceTail2FullInst = nullptr;
    ShadowFn shadowHandInst = nullptr, shadowTailInst = nullptr, shadowTail2Inst = nullptr;
    // depth 0 with the first-bounce sort fused in (kFirstTile-thread workgroups)
    FirstFn bounceFirst = nullptr, bounceFirstFull = nullptr;
    // the hand-off in batches (BVH8): pass kernels that hand off and their tails; rolling steps [full][per-lane]
    JobsFn bounceJobsHand = nullptr, bounceJobsHandFull = nullptr, bounceJobsTail = nullptr, bounceJobsTailFull = nullptr;
    JobsFn shadowJobsHand = nullptr, shadowJobsTail = nullptr;
    RollFn bounceRollHand[2][2] = {}, bounceRollTail[2][2] = {}, bounceRollTail2[2][2] = {};   // tail2: second level
    RollFn shadowRollHand = nullptr, shadowRollTail = nullptr;
};

template <int V, int TRI, int SL>
void pickSL(int waves, Kernels& k) {
    k.bounce = waves >= 8 ? k_bounce<V, TRI, SL, 8, false> : (waves >= 6 ? k_bounce<V, TRI, SL, 6, false> : k_bounce<V, TRI, SL, 1, false>);
    k.tracePaths = k_trace_paths<V, TRI, SL, kRefill>;
    k.tracePlain = k_trace_paths<V, TRI, SL, kWave>;
    k.shadow = k_shadow<V, TRI, SL, false>;
    k.shadowRefill = k_shadow_refill<V, TRI, SL, kRefill>;
    k.traceC = k_trace<V, TRI, false, SL, false>;
    k.traceA = k_trace<V, TRI, true, SL, false>;
    k.bounceInst = k_bounce<V, TRI, SL, kInstWaves, true>;
    k.shadowInst = k_shadow<V, TRI, SL, true>;
    k.traceCInst = k_trace<V, TRI, false, SL, true>;
    k.traceAInst = k_trace<V, TRI, true, SL, true>;
    k.bounceJobs = k_bounce_jobs<V, TRI, SL, 8>;
    k.shadowJobs = k_shadow_jobs<V, TRI, SL>;
    k.bounceFull = k_bounce<V, TRI, SL, kFullWaves, false, true>;
    k.bounceFullInst = k_bounce<V, TRI, SL, kFullWaves, true, true>;
    k.bounceRefill = waves >= 8 ? k_bounce_refill<V, TRI, SL, 8, false, kShadeBatch>
                                : (waves >= 6 ? k_bounce_refill<V, TRI, SL, 6, false, kShadeBatch> : k_bounce_refill<V, TRI, SL, 4, false, kShadeBatch>);
    k.bounceRefillFull = k_bounce_refill<V, TRI, SL, kFullWaves, true, kShadeBatch>;
    k.bounceJobsFull = k_bounce_jobs<V, TRI, SL, kFullWaves, true>;
    k.bounceRoll = k_bounce_roll<V, TRI, SL, 8>;
    k.bounceRollFull = k_bounce_roll<V, TRI, SL, kFullWaves, true>;
    k.bounceRollLane = k_bounce_roll<V, TRI, SL, 8, false, true>;
    k.bounceRollFullLane = k_bounce_roll<V, TRI, SL, kFullWaves, true, true>;
    k.shadowRoll = k_shadow_roll<V, TRI, SL>;
    if constexpr (V == 8) {
        k.bounceHand = k_bounce<V, TRI, SL, 8, false, false, true>;
        k.bounceTail = k_bounce_tail<TRI, SL, 6, false>;
        k.bounceHandFull = k_bounce<V, TRI, SL, kFullWaves, false, true, true>;
        k.bounceTailFull = k_bounce_tail<TRI, SL, kFullWaves, true>;
        k.shadowHand = k_shadow<V, TRI, SL, false, true>;
        k.shadowTail = k_shadow_tail<TRI, SL>;
        k.bounceTail2 = k_bounce_tail<TRI, SL, 8, false, 2>;
        k.bounceTail2Full = k_bounce_tail<TRI, SL, kFullWaves, true, 2>;
        k.shadowTail2 = k_shadow_tail<TRI, SL, 2>;
        k.bounceHandInst = k_bounce<V, TRI, SL, kInstWaves, true, false, true>;
        k.bounceHandFullInst = k_bounce<V, TRI, SL, kFullWaves, true, true, true>;
        k.bounceTailInst = k_bounce_tail<TRI, SL, kInstWaves, false, 1, true>;
        k.bounceTailFullInst = k_bounce_tail<TRI, SL, kFullWaves, true, 1, true>;
        k.bounceTail2Inst = k_bounce_tail<TRI, SL, kInstWaves, false, 2, true>;
        k.bounceTail2FullInst = k_bounce_tail<TRI, SL, kFullWaves, true, 2, true>;
        k.shadowHandInst = k_shadow<V, TRI, SL, true, true>;
        k.shadowTailInst = k_shadow_tail<TRI, SL, 1, true>;
        k.shadowTail2Inst = k_shadow_tail<TRI, SL, 2, true>;
        k.bounceJobsHand = k_bounce_jobs<V, TRI, SL, 8, false, true>;
        k.bounceJobsHandFull = k_bounce_jobs<V, TRI, SL, kFullWaves, true, true>;
        k.bounceJobsTail = k_bounce_jobs_tail<TRI, SL, 8, false>;
        k.bounceJobsTailFull = k_bounce_jobs_tail<TRI, SL, kFullWaves, true>;
        k.shadowJobsHand = k_shadow_jobs<V, TRI, SL, true>;
        k.shadowJobsTail = k_shadow_jobs_tail<TRI, SL, false, DevJobTable>;
        k.bounceRollHand[0][0] = k_bounce_roll<V, TRI, SL, 8, false, false, true>;
        k.bounceRollHand[0][1] = k_bounce_roll<V, TRI, SL, 8, false, true, true>;
        k.bounceRollHand[1][0] = k_bounce_roll<V, TRI, SL, kFullWaves, true, false, true>;
        k.bounceRollHand[1][1] = k_bounce_roll<V, TRI, SL, kFullWaves, true, true, true>;
        k.bounceRollTail[0][0] = k_bounce_roll_tail<TRI, SL, 8, false, false>;
        k.bounceRollTail[0][1] = k_bounce_roll_tail<TRI, SL, 8, false, true>;
        k.bounceRollTail[1][0] = k_bounce_roll_tail<TRI, SL, kFullWaves, true, false>;
        k.bounceRollTail[1][1] = k_bounce_roll_tail<TRI, SL, kFullWaves, true, true>;
        k.bounceRollTail2[0][0] = k_bounce_roll_tail<TRI, SL, 8, false, false, 2>;
        k.bounceRollTail2[0][1] = k_bounce_roll_tail<TRI, SL, 8, false, true, 2>;
        k.bounceRollTail2[1][0] = k_bounce_roll_tail<TRI, SL, kFullWaves, true, false, 2>;
        k.bounceRollTail2[1][1] = k_bounce_roll_tail<TRI, SL, kFullWaves, true, true, 2>;
        k.shadowRollHand = k_shadow_roll<V, TRI, SL, true>;
        k.shadowRollTail = k_shadow_jobs_tail<TRI, SL, true, DevStepTable>;
    }
    k.bounceFirst = k_bounce_first<V, TRI, SL, 6, false>;
    k.bounceFirstFull = k_bounce_first<V, TRI, SL, kFullWaves, true>;
}
template <int V, int TRI>
void pickTri(int sl, int waves, Kernels& k) {
    if (sl == 8) pickSL<V, TRI, 8>(waves, k);
    else if (sl == 32) pickSL<V, TRI, 32>(waves, k);
    else pickSL<V, TRI, 16>(waves, k);
}
// the kernels of layout V for a triangle test, LDS stack size and bounce register budget (explicitly instantiated once
// per layout in kernels_v<V>.hip)
template <int V>
void pickKernels(int tri, int sl, int waves, Kernels& k) {
    if (tri == kTriMT) pickTri<V, kTriMT>(sl, waves, k);
    else if (tri == kTriWatertight) pickTri<V, kTriWatertight>(sl, waves, k);
    else pickTri<V, kTriWoop>(sl, waves, k);
}

// layout-independent kernels (device/common.hip)
__global__ void k_shade(DevScene S, DevPaths P, const FrameParams* __restrict__ Fp, int depth, int smpBase, int K);
__global__ void k_pass_begin(int* cnt, int words, long long NK);
// continuation sort of the depth-(depth + 1) queue (P.qs, P.qk) into P.q[(depth + 1) & 1], per partition: bucket
// sizes, then the scatter. Grids of kSortBlock-thread workgroups, a multiple of kParts.
constexpr int kSortBlock = 256, kSortPer = 16, kSortTile = kSortBlock * kSortPer;
__global__ void k_qcount(DevPaths P, int depth);
__global__ void k_qscatter(DevPaths P, int depth);
__global__ void k_qsort_first(DevPaths P);
__global__ void k_pass_end(const int* cnt, int depths, unsigned long long* tot);
__global__ void k_reduce(const Vec4* __restrict__ L, Vec4* __restrict__ accum, int N, int K);
__global__ void k_fill(uint32_t* p, size_t words, uint32_t v);
// per triangle of a scene's corner array: (cross(e1, e2), material id bits of the second corner)
__global__ void k_tri_geo(const Vec4* __restrict__ verts, int nt, Vec4* __restrict__ geo);
__global__ void k_pass_begin_jobs(int* cnt, int words, const DevJobTable* T);
__global__ void k_reduce_jobs(const DevJobTable* __restrict__ T, const Vec4* __restrict__ L);
// rolling batches: the step's counter block (zeroed, then each partition's input = the previous step's continuations
// + its share of the camera rays), the reduction of the jobs done after the step, the step's totals
__global__ void k_step_begin(const DevStepTable* T, int* cnt, int step);
__global__ void k_reduce_roll(const DevStepTable* __restrict__ T, const Vec4* __restrict__ L);
__global__ void k_step_end(const int* cnt, int step, unsigned long long* tot);

extern template void pickKernels<2>(int, int, int, Kernels&);
extern template void pickKernels<4>(int, int, int, Kernels&);
extern template void pickKernels<8>(int, int, int, Kernels&);

}  // namespace dev
}  // namespace cwp
