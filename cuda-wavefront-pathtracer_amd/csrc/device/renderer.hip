// device/renderer.hip - gfx950 wavefront path tracer: fused raygen/trace/shade bounce kernel, shadow-ray kernel,
// per-pixel reduction, and the host driver that enqueues a whole frame without host synchronisation.
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
#include "core/sampling.h"
#include "device/renderer.h"
#include "device/traverse.h"
#include "host/scene_check.h"
#include <hip/hip_runtime.h>
#include <algorithm>
#include <cstdio>

namespace cwp {

namespace {

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

struct DevScene {
    const Vec4* nodes;    // BVH layout (BVH8: 128-byte node stride)
    const Vec4* tri;      // traversal triangle data (Woop rows or corners)
    const Vec4* verts;    // corners for shading
    const Vec4* nrm;      // per-corner normals or null
    const Vec4* albedo;   // per material
    int root;
};

// Path state. Path id i = pixel * K + sample: the 64 lanes of a wave start as 64 samples of one pixel, and each
// partition's slice of the id space is a band of the image.
struct DevPaths {
    Vec4 *O, *D, *T, *L;  // origin, direction, throughput, radiance per path id
    Vec4* H;              // closest hit of the current bounce (t, u, v, triangle bits) when tracing and shading are split
    int* q[2];            // ping-pong queues of path ids: depth d reads q[d & 1]
    Vec4 *sO, *sD, *sC;   // shadow rays of the current bounce: (origin, path id bits), (dir, tmax), (contribution)
    int* cnt;             // kCntStride counters per depth (one line per partition)
    void* spill;          // traversal-stack spill slab, spillCap entries per thread
    int spillCap;
};

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

// start of partition `part` in a region laid out as the partitions' queue lengths of a depth's counter block
__device__ __forceinline__ int partStart(const int* cnt, int part) {
    int s = 0;
    for (int y = 0; y < part; y++) s += ctrv(cnt, y, kCQueue);
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
struct ShadeOut { bool cont, shadow, miss; Vec3 po, nd, Tn, sdir, scon; };

__device__ __forceinline__ ShadeOut shadeLegacy(const DevScene& S, const DevPaths& P, const FrameParams& F, int depth,
                                                int i, int pix, uint32_t smp, Vec3 o, Vec3 d, Vec3 T, const Hit& h) {
    ShadeOut out;
    out.cont = false; out.shadow = false; out.miss = false;
    out.po = v3(0.f, 0.f, 0.f); out.nd = out.po; out.Tn = out.po; out.sdir = out.po; out.scon = out.po;
    if (h.tri < 0) {
        out.miss = true;
        Vec3 c = T * F.sky;
        if (depth == 0) P.L[i] = v4(c, 0.f);
        else { Vec4 l = P.L[i]; l.x += c.x; l.y += c.y; l.z += c.z; P.L[i] = l; }
        return out;
    }
    if (depth == 0) P.L[i] = v4(0.f, 0.f, 0.f, 0.f);
    if (depth > F.maxB) return out;
    const Vec3 p = o + d * h.t;
    const Vec4 a = S.verts[3 * h.tri], b = S.verts[3 * h.tri + 1], c = S.verts[3 * h.tri + 2];
    const Vec3 e1 = xyz(b) - xyz(a), e2 = xyz(c) - xyz(a);
    Vec3 ng = normalize(cross(e1, e2)), ns = ng;
    if (F.hasN) {
        Vec3 n0 = xyz(S.nrm[3 * h.tri]), n1 = xyz(S.nrm[3 * h.tri + 1]), n2 = xyz(S.nrm[3 * h.tri + 2]);
        Vec3 nn = n0 * (1.f - h.u - h.v) + n1 * h.u + n2 * h.v;
        float l = length(nn);
        if (l > 0.f) ns = nn * (1.f / l);
        if (dot(ns, ng) < 0.f) ns = -ns;
    }
    if (dot(ng, d) > 0.f) { ng = -ng; ns = -ns; }
    const Vec3 alb = xyz(S.albedo[asInt(b.w)]);
    out.po = p + ng * F.eps;
    const uint32_t dd = (uint32_t)depth + 1u;
    const Vec3 l = uniformCone(F.sunTo, F.sunCos, rnd((uint32_t)pix, smp, dd, 2, F.seed), rnd((uint32_t)pix, smp, dd, 3, F.seed));
    const float cl = dot(ns, l);
    if (cl > 0.f && dot(ng, l) > 0.f) { out.shadow = true; out.sdir = l; out.scon = T * alb * F.sunE * (cl * kInvPi); }
    out.nd = cosineHemisphere(ns, rnd((uint32_t)pix, smp, dd, 0, F.seed), rnd((uint32_t)pix, smp, dd, 1, F.seed));
    if (dot(out.nd, ng) > 0.f) { out.cont = true; out.Tn = T * alb; }
    return out;
}

// Wave-level appends of the shading results: continuations to partition `part` of the next queue, shadow rays to
// partition `part` of the shadow queue. Returns the wave's number of misses.
__device__ __forceinline__ int appendResults(const DevPaths& P, int* cnt, int part, int outStart, int* qOut, int i,
                                             const ShadeOut& r) {
    const uint64_t mc = __ballot(r.cont), ms = __ballot(r.shadow), mm = __ballot(r.miss);
    if (mc) {
        const int pos = outStart + waveAppend(mc, ctr(cnt + kCntStride, part, kCQueue));
        if (r.cont) { qOut[pos] = i; P.O[i] = v4(r.po, 0.f); P.D[i] = v4(r.nd, 0.f); P.T[i] = v4(r.Tn, 0.f); }
    }
    if (ms) {
        const int pos = outStart + waveAppend(ms, ctr(cnt, part, kCShadow));
        if (r.shadow) { P.sO[pos] = v4(r.po, asFloat(i)); P.sD[pos] = v4(r.sdir, 1e30f); P.sC[pos] = v4(r.scon, 0.f); }
    }
    return (int)__builtin_popcountll(mm);
}

// The ray of path i at this depth: its camera ray at depth 0 (recomputed from the path id), else the stored state.
__device__ __forceinline__ void pathRay(const DevPaths& P, const FrameParams& F, int depth, int i, int smpBase, int K,
                                        Vec3& o, Vec3& d, float& tmin) {
    if (depth == 0) { cameraRay(F, i / K, (uint32_t)(smpBase + i % K), o, d); tmin = F.nearT; }
    else { o = xyz(P.O[i]); d = xyz(P.D[i]); tmin = 0.f; }
}

// Queue position of the f-th entry a wave hands out: the wave walks the chunks lw, lw + nw, lw + 2 nw, ... of its
// partition, 64 entries each (increasing in f).
__device__ __forceinline__ int fetchPos(uint32_t f, int lw, int nw) { return (lw + (int)(f >> 6) * nw) * kWave + (int)(f & 63u); }

// idle lanes of a persistent wave are refilled once at least this many are idle (kWave: only when all are, which is
// the plain one-ray-per-lane schedule)
constexpr int kRefill = 16;

// One bounce of the legacy integrator over the queue of depth `depth` (depth 0: every path id, camera rays made here),
// traversal and shading fused. The queue is split into kParts partitions, one per XCD; the waves of XCD x walk
// partition x, and append their continuations to partition x of the next queue and their shadow rays to partition x of
// the shadow queue, so a path stays on one XCD (and in its L2) for its whole life. Partition x of the depth-(d+1) queue
// starts where the depth-d partitions before x end, which bounds it by the rays partition x traced.
// WAVES: minimum waves per SIMD the register allocation must allow (1 = unconstrained).
template <int V, int TRI, int SL, int WAVES>
__global__ void __launch_bounds__(kBlock, WAVES) k_bounce(DevScene S, DevPaths P, FrameParams F, int depth, int smpBase, int K) {
    using E = typename StackEntry<V>::type;
    __shared__ E lds[SL > 0 ? SL * kBlock : 1];
    const int lane = laneId();
    const int gtid = blockIdx.x * kBlock + threadIdx.x;
    LaneStack<E, SL> st{ (LdsPtr<E>)(lds + threadIdx.x), (E*)P.spill + (size_t)gtid * P.spillCap, 0 };
    const int part = (int)(blockIdx.x % kParts), lw = (int)(blockIdx.x / kParts), nw = (int)(gridDim.x / kParts);
    int* cnt = P.cnt + depth * kCntStride;
    const int n = ctrv(cnt, part, kCQueue);
    const int inStart = depth == 0 ? partStart(cnt, part) : partStart(cnt - kCntStride, part);
    const int outStart = partStart(cnt, part);
    const int* qIn = P.q[depth & 1];
    int* qOut = P.q[(depth + 1) & 1];
    int misses = 0;
    for (int base = lw * kWave; base < n; base += nw * kWave) {
        const int k = base + lane;
        int i = 0;
        ShadeOut res;
        res.cont = false; res.shadow = false; res.miss = false;
        if (k < n) {
            i = depth == 0 ? inStart + k : qIn[inStart + k];
            Vec3 o, d; float tmin;
            pathRay(P, F, depth, i, smpBase, K, o, d, tmin);
            const RayPre r = prepareRay<TRI>(o, d, tmin);
            Hit h; h.t = 1e30f; h.u = 0.f; h.v = 0.f; h.tri = -1;
            st.sp = 0;
            traceScene<V, false, TRI>(S, r, h, st);
            // throughput is fetched after the traversal so that no registers hold it across the traversal loop
            const Vec3 T = depth == 0 ? v3(1.f, 1.f, 1.f) : xyz(P.T[i]);
            res = shadeLegacy(S, P, F, depth, i, i / K, (uint32_t)(smpBase + i % K), r.o, r.d, T, h);
        }
        misses += appendResults(P, cnt, part, outStart, qOut, i, res);
    }
    waveAdd(misses, ctr(cnt, part, kCMiss));
}

// Closest hits of the depth-`depth` queue with persistent lanes and dynamic ray fetching (after Aila & Laine 2009;
// reference: persistentTrace, src/render/wavefront.cu:403-434): a lane whose ray finishes stays idle until at least
// REFILL lanes of its wave are idle, then they all take the next entries of the wave's share of its partition, so
// long rays do not leave the rest of the wave empty. Hits go to H[path id] for k_shade.
template <int V, int TRI, int SL, int REFILL>
__global__ void __launch_bounds__(kBlock, 8) k_trace_paths(DevScene S, DevPaths P, FrameParams F, int depth, int smpBase, int K) {
    using E = typename StackEntry<V>::type;
    __shared__ E lds[SL > 0 ? SL * kBlock : 1];
    const int gtid = blockIdx.x * kBlock + threadIdx.x;
    LaneStack<E, SL> st{ (LdsPtr<E>)(lds + threadIdx.x), (E*)P.spill + (size_t)gtid * P.spillCap, 0 };
    const int part = (int)(blockIdx.x % kParts), lw = (int)(blockIdx.x / kParts), nw = (int)(gridDim.x / kParts);
    const int* cnt = P.cnt + depth * kCntStride;
    const int n = ctrv(cnt, part, kCQueue);
    const int inStart = depth == 0 ? partStart(cnt, part) : partStart(cnt - kCntStride, part);
    const int* qIn = P.q[depth & 1];
    int i = -1;                       // path id of this lane's ray, -1 = idle
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
                    i = depth == 0 ? inStart + k : qIn[inStart + k];
                    Vec3 o, d; float tmin;
                    pathRay(P, F, depth, i, smpBase, K, o, d, tmin);
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

// Shading of the depth-`depth` queue from the hits of k_trace_paths (no traversal: every lane of a wave shades).
__global__ void __launch_bounds__(kBlock) k_shade(DevScene S, DevPaths P, FrameParams F, int depth, int smpBase, int K) {
    const int lane = laneId();
    const int part = (int)(blockIdx.x % kParts), lw = (int)(blockIdx.x / kParts), nw = (int)(gridDim.x / kParts);
    int* cnt = P.cnt + depth * kCntStride;
    const int n = ctrv(cnt, part, kCQueue);
    const int inStart = depth == 0 ? partStart(cnt, part) : partStart(cnt - kCntStride, part);
    const int outStart = partStart(cnt, part);
    const int* qIn = P.q[depth & 1];
    int* qOut = P.q[(depth + 1) & 1];
    int misses = 0;
    for (int base = lw * kWave; base < n; base += nw * kWave) {
        const int k = base + lane;
        int i = 0;
        ShadeOut res;
        res.cont = false; res.shadow = false; res.miss = false;
        if (k < n) {
            i = depth == 0 ? inStart + k : qIn[inStart + k];
            Vec3 o, d; float tmin;
            pathRay(P, F, depth, i, smpBase, K, o, d, tmin);
            const Vec4 hv = P.H[i];
            Hit h; h.t = hv.x; h.u = hv.y; h.v = hv.z; h.tri = asInt(hv.w);
            const Vec3 T = depth == 0 ? v3(1.f, 1.f, 1.f) : xyz(P.T[i]);
            res = shadeLegacy(S, P, F, depth, i, i / K, (uint32_t)(smpBase + i % K), o, d, T, h);
        }
        misses += appendResults(P, cnt, part, outStart, qOut, i, res);
    }
    waveAdd(misses, ctr(cnt, part, kCMiss));
}

// Shadow rays of bounce `depth` (partition per XCD as in k_bounce): any hit; unoccluded contributions are added to
// the path's radiance. The any-hit traversal fits 64 registers without spilling: 8 waves per SIMD.
template <int V, int TRI, int SL>
__global__ void __launch_bounds__(kBlock, 8) k_shadow(DevScene S, DevPaths P, int depth) {
    using E = typename StackEntry<V>::type;
    __shared__ E lds[SL > 0 ? SL * kBlock : 1];
    const int gtid = blockIdx.x * kBlock + threadIdx.x;
    LaneStack<E, SL> st{ (LdsPtr<E>)(lds + threadIdx.x), (E*)P.spill + (size_t)gtid * P.spillCap, 0 };
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
            const RayPre r = prepareRay<TRI>(xyz(O), xyz(D), 0.f);
            Hit h; h.t = D.w; h.u = 0.f; h.v = 0.f; h.tri = -1;
            st.sp = 0;
            if (!traceScene<V, true, TRI>(S, r, h, st)) {
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

// Shadow rays with persistent lanes and dynamic ray fetching (as k_trace_paths).
template <int V, int TRI, int SL, int REFILL>
__global__ void __launch_bounds__(kBlock, 8) k_shadow_refill(DevScene S, DevPaths P, int depth) {
    using E = typename StackEntry<V>::type;
    __shared__ E lds[SL > 0 ? SL * kBlock : 1];
    const int gtid = blockIdx.x * kBlock + threadIdx.x;
    LaneStack<E, SL> st{ (LdsPtr<E>)(lds + threadIdx.x), (E*)P.spill + (size_t)gtid * P.spillCap, 0 };
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

// Closest or any hit of explicit rays (parity and kernel-level entry point).
template <int V, int TRI, bool ANY, int SL>
__global__ void __launch_bounds__(kBlock, 8) k_trace(DevScene S, const Vec4* __restrict__ rays, int n, Vec4* __restrict__ hits, void* spill, int spillCap) {
    using E = typename StackEntry<V>::type;
    __shared__ E lds[SL > 0 ? SL * kBlock : 1];
    const int gtid = blockIdx.x * kBlock + threadIdx.x;
    LaneStack<E, SL> st{ (LdsPtr<E>)(lds + threadIdx.x), (E*)spill + (size_t)gtid * spillCap, 0 };
    for (int k = gtid; k - (int)threadIdx.x < n; k += gridDim.x * kBlock) {
        if (k >= n) continue;
        const Vec4 O = rays[2 * (size_t)k], D = rays[2 * (size_t)k + 1];
        const RayPre r = prepareRay<TRI>(xyz(O), xyz(D), O.w);
        Hit h; h.t = D.w; h.u = 0.f; h.v = 0.f; h.tri = -1;
        st.sp = 0;
        traceScene<V, ANY, TRI>(S, r, h, st);
        hits[k] = v4(h.t, h.u, h.v, asFloat(h.tri));
    }
}

// zero the pass counters; the depth-0 queue of partition x is the x-th of kParts equal slices of the path ids
__global__ void k_pass_begin(int* cnt, int words, long long NK) {
    for (int j = threadIdx.x; j < words; j += blockDim.x) cnt[j] = 0;
    __syncthreads();
    if (threadIdx.x < kParts) {
        const long long x = threadIdx.x;
        *ctr(cnt, (int)x, kCQueue) = (int)((x + 1) * NK / kParts - x * NK / kParts);
    }
}

// per-pass totals: [0] closest rays, [1] shadow rays, [2] misses, [3] unoccluded, [4] continuations,
// [5] camera-ray misses
__global__ void k_pass_end(const int* cnt, int depths, unsigned long long* tot) {
    if (threadIdx.x != 0) return;
    unsigned long long a = 0, b = 0, m = 0, u = 0, c = 0, m0 = 0;
    for (int x = 0; x < kParts; x++) m0 += (unsigned)ctrv(cnt, x, kCMiss);
    for (int d = 0; d < depths; d++) {
        const int* cd = cnt + d * kCntStride;
        unsigned long long q = 0;
        for (int x = 0; x < kParts; x++) {
            q += (unsigned)ctrv(cd, x, kCQueue); b += (unsigned)ctrv(cd, x, kCShadow);
            m += (unsigned)ctrv(cd, x, kCMiss); u += (unsigned)ctrv(cd, x, kCUnocc);
        }
        a += q; if (d > 0) c += q;
    }
    tot[0] += a; tot[1] += b; tot[2] += m; tot[3] += u; tot[4] += c; tot[5] += m0;
}

// accum[pix] += sum over the K samples of the pass (path ids pix*K .. pix*K + K-1), in sample order. One wave per 64
// pixels: the K samples of a pixel are contiguous, so the wave loads chunks of 16 samples of 4 pixels per instruction
// (4 runs of 256 B) into LDS, then each lane adds its pixel's samples in order.
constexpr int kRedChunk = 16;
__global__ void __launch_bounds__(kWave) k_reduce(const Vec4* __restrict__ L, Vec4* __restrict__ accum, int N, int K) {
    __shared__ float buf[kWave][3 * kRedChunk + 1];   // +1: rows 49 words apart, no bank conflicts between lanes
    const int lane = laneId();
    for (int p0 = blockIdx.x * kWave; p0 < N; p0 += gridDim.x * kWave) {
        float r = 0.f, g = 0.f, b = 0.f;
        for (int s0 = 0; s0 < K; s0 += kRedChunk) {
            const int cs = min(kRedChunk, K - s0);
            for (int j = 0; j < kWave / 4; j++) {
                const int p = 4 * j + (lane >> 4), s = lane & 15;
                if (p0 + p < N && s < cs) {
                    const Vec4 v = L[(size_t)(p0 + p) * K + s0 + s];
                    buf[p][3 * s] = v.x; buf[p][3 * s + 1] = v.y; buf[p][3 * s + 2] = v.z;
                }
            }
            __syncthreads();
            for (int s = 0; s < cs; s++) { r += buf[lane][3 * s]; g += buf[lane][3 * s + 1]; b += buf[lane][3 * s + 2]; }
            __syncthreads();
        }
        const int pix = p0 + lane;
        if (pix < N) { Vec4 a = accum[pix]; a.x += r; a.y += g; a.z += b; accum[pix] = a; }
    }
}

__global__ void k_fill(uint32_t* p, size_t words, uint32_t v) {
    for (size_t j = blockIdx.x * (size_t)blockDim.x + threadIdx.x; j < words; j += (size_t)gridDim.x * blockDim.x) p[j] = v;
}

// ------------------------------------------------------------------------------------------------------- dispatch
using BounceFn = void (*)(DevScene, DevPaths, FrameParams, int, int, int);
using ShadowFn = void (*)(DevScene, DevPaths, int);
using TraceFn = void (*)(DevScene, const Vec4*, int, Vec4*, void*, int);

// the kernels of one (layout, triangle test, LDS stack) configuration
struct Kernels {
    BounceFn bounce = nullptr;                          // fused trace + shade
    BounceFn tracePaths = nullptr, tracePlain = nullptr; // split: closest-hit trace with / without refill (then k_shade)
    ShadowFn shadow = nullptr, shadowRefill = nullptr;
    TraceFn traceC = nullptr, traceA = nullptr;
};

template <int V, int TRI, int SL>
void pickSL(int waves, Kernels& k) {
    k.bounce = waves >= 8 ? k_bounce<V, TRI, SL, 8> : (waves >= 6 ? k_bounce<V, TRI, SL, 6> : k_bounce<V, TRI, SL, 1>);
    k.tracePaths = k_trace_paths<V, TRI, SL, kRefill>;
    k.tracePlain = k_trace_paths<V, TRI, SL, kWave>;
    k.shadow = k_shadow<V, TRI, SL>;
    k.shadowRefill = k_shadow_refill<V, TRI, SL, kRefill>;
    k.traceC = k_trace<V, TRI, false, SL>;
    k.traceA = k_trace<V, TRI, true, SL>;
}
template <int V, int TRI>
void pickTri(int sl, int waves, Kernels& k) {
    if (sl == 8) pickSL<V, TRI, 8>(waves, k);
    else if (sl == 32) pickSL<V, TRI, 32>(waves, k);
    else pickSL<V, TRI, 16>(waves, k);
}
template <int V>
void pickKernels(int tri, int sl, int waves, Kernels& k) {
    if (tri == kTriMT) pickTri<V, kTriMT>(sl, waves, k);
    else if (tri == kTriWatertight) pickTri<V, kTriWatertight>(sl, waves, k);
    else pickTri<V, kTriWoop>(sl, waves, k);
}

#define HIPCHK(x)                                                                                                      \
    do {                                                                                                               \
        hipError_t e_ = (x);                                                                                           \
        if (e_ != hipSuccess) { err = std::string(#x) + ": " + hipGetErrorString(e_); return false; }                  \
    } while (0)

}  // namespace

struct Renderer::Impl {
    RenderOptions O;
    FrameParams F;
    int N = 0, K = 1, NK = 0, depths = 0, sl = 16;
    hipStream_t st = nullptr;
    hipEvent_t f0 = nullptr, f1 = nullptr;
    std::vector<hipEvent_t> evs; size_t evUsed = 0;
    DevScene S{};
    DevPaths P{};
    Vec4* accum = nullptr;
    unsigned long long* tot = nullptr;
    unsigned long long* hostTot = nullptr;
    std::vector<void*> allocs;
    Kernels kn;
    BounceFn bounce = nullptr, shade = nullptr;   // fused: bounce; split: bounce = trace kernel, then shade
    ShadowFn shadow = nullptr;
    int gridBounce = 0, gridShade = 0, gridShadow = 0, gridTrace = 0, maxGrid = 0;
    void* stageRays = nullptr; void* stageHits = nullptr; size_t stageCap = 0;   // traceHost staging buffers

    template <class T> bool alloc(T*& p, size_t count, std::string& err) {
        void* q = nullptr;
        size_t bytes = std::max<size_t>(count * sizeof(T), 16);
        HIPCHK(hipMalloc(&q, bytes));
        allocs.push_back(q); p = (T*)q;
        return true;
    }
    template <class T> bool upload(T*& p, const std::vector<T>& v, std::string& err) {
        if (v.empty()) { p = nullptr; return true; }
        if (!alloc(p, v.size(), err)) return false;
        HIPCHK(hipMemcpy(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
        return true;
    }
    hipEvent_t nextEvent() {
        if (evUsed == evs.size()) { hipEvent_t e; if (hipEventCreate(&e) != hipSuccess) return nullptr; evs.push_back(e); }
        return evs[evUsed++];
    }
    bool fill(void* p, size_t bytes, std::string& err) {
        size_t words = bytes / 4;
        int g = (int)std::min<size_t>((words + 255) / 256, 4096);
        if (g < 1) g = 1;
        k_fill<<<g, 256, 0, st>>>((uint32_t*)p, words, 0u);
        HIPCHK(hipGetLastError());
        return true;
    }
};

Renderer::Renderer() : p_(new Impl) {}
Renderer::~Renderer() {
    if (!p_) return;
    if (p_->st) hipStreamSynchronize(p_->st);
    for (void* a : p_->allocs) hipFree(a);
    if (p_->stageRays) hipFree(p_->stageRays);
    if (p_->stageHits) hipFree(p_->stageHits);
    for (hipEvent_t e : p_->evs) hipEventDestroy(e);
    if (p_->f0) hipEventDestroy(p_->f0);
    if (p_->f1) hipEventDestroy(p_->f1);
    if (p_->hostTot) hipHostFree(p_->hostTot);
    if (p_->st) hipStreamDestroy(p_->st);
    delete p_;
}
const RenderOptions& Renderer::options() const { return p_->O; }

bool deviceInfo(int& numCU, std::string& name, std::string& err) {
    int n = 0;
    HIPCHK(hipGetDeviceCount(&n));
    if (n < 1) { err = "no HIP device"; return false; }
    hipDeviceProp_t prop;
    HIPCHK(hipGetDeviceProperties(&prop, 0));
    numCU = prop.multiProcessorCount; name = prop.name;
    return true;
}

bool deviceMalloc(size_t bytes, void** ptr, std::string& err) {
    *ptr = nullptr;
    HIPCHK(hipMalloc(ptr, std::max<size_t>(bytes, 16)));
    return true;
}
void deviceFree(void* ptr) { if (ptr) hipFree(ptr); }
bool deviceCopy(void* dst, const void* src, size_t bytes, int kind, std::string& err) {
    if (!bytes) return true;
    const hipMemcpyKind k = kind == 1 ? hipMemcpyHostToDevice : (kind == 2 ? hipMemcpyDeviceToHost : hipMemcpyDeviceToDevice);
    HIPCHK(hipMemcpy(dst, src, bytes, k));
    return true;
}

bool Renderer::init(const SceneUpload& S, const FrameParams& F, const RenderOptions& O, std::string& err) {
    Impl& I = *p_;
    // the kernels trust every index in the scene: it is checked on the host first
    int stackNeed = 0;
    if (!checkScene(S, stackNeed, err)) return false;
    if (O.tri < kTriMT || O.tri > kTriWatertight) { err = "unknown triangle test"; return false; }
    if (O.tri == kTriWoop && S.woop.size() != S.verts.size()) { err = "Woop rows missing"; return false; }
    if (F.cam.w <= 0 || F.cam.h <= 0) { err = "image size must be positive"; return false; }
    if (O.stackLds != 8 && O.stackLds != 16 && O.stackLds != 32) { err = "stack_lds must be 8, 16 or 32"; return false; }
    if (O.bounceWaves != 1 && O.bounceWaves != 6 && O.bounceWaves != 8) { err = "bounce_waves must be 1, 6 or 8"; return false; }
    if (O.bounces < 0 || O.bounces > 64) { err = "bounces must be in [0, 64]"; return false; }
    I.O = O; I.F = F;
    I.sl = O.stackLds == 8 ? 8 : (O.stackLds == 32 ? 32 : 16);
    I.O.stackLds = I.sl;
    HIPCHK(hipSetDevice(0));
    if (!deviceInfo(numCU, gpuName, err)) return false;
    size_t free0 = 0, total = 0;
    HIPCHK(hipMemGetInfo(&free0, &total));
    HIPCHK(hipStreamCreateWithFlags(&I.st, hipStreamNonBlocking));
    HIPCHK(hipEventCreate(&I.f0)); HIPCHK(hipEventCreate(&I.f1));

    // BVH8 nodes are re-strided from 80 to 128 bytes: each node then fills exactly one 128-byte cache line instead of
    // straddling two in 3 cases out of 5
    std::vector<Vec4> devNodes;
    if (S.variant == 8) {
        const size_t nn = S.nodes.size() / 5;
        devNodes.assign(nn * kBvh8Stride, v4(0.f, 0.f, 0.f, 0.f));
        for (size_t j = 0; j < nn; j++)
            for (int q = 0; q < 5; q++) devNodes[j * kBvh8Stride + q] = S.nodes[j * 5 + q];
    } else devNodes = S.nodes;
    Vec4 *dn, *dw, *dv, *dnr, *da;
    if (!I.upload(dn, devNodes, err) || !I.upload(dw, S.woop, err) || !I.upload(dv, S.verts, err) ||
        !I.upload(dnr, S.normals, err) || !I.upload(da, S.albedo, err)) return false;
    I.S.nodes = dn; I.S.verts = dv; I.S.nrm = dnr; I.S.albedo = da; I.S.root = S.root;
    I.S.tri = O.tri == kTriWoop ? dw : dv;
    if (F.hasN && !dnr) { err = "hasN set without normals"; return false; }

    if (O.fuse != 0 && O.fuse != 1) { err = "fuse must be 0 or 1"; return false; }
    if (O.persistent && O.fuse) { err = "persistent traversal needs the split kernels (fuse 0)"; return false; }
    if (S.variant == 2) pickKernels<2>(O.tri, I.sl, O.bounceWaves, I.kn);
    else if (S.variant == 4) pickKernels<4>(O.tri, I.sl, O.bounceWaves, I.kn);
    else pickKernels<8>(O.tri, I.sl, O.bounceWaves, I.kn);
    I.bounce = O.fuse ? I.kn.bounce : (O.persistent ? I.kn.tracePaths : I.kn.tracePlain);
    I.shade = O.fuse ? nullptr : k_shade;
    I.shadow = (O.persistent || O.shadowRefill) ? I.kn.shadowRefill : I.kn.shadow;
    auto gridOf = [&](const void* fn) {
        int b = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, fn, kBlock, 0) != hipSuccess || b < 1) b = 1;
        if (O.blocksPerCU > 0) b = std::min(b, O.blocksPerCU);
        return std::max(kParts, (b * numCU) / kParts * kParts);   // a whole number of waves per partition
    };
    I.gridBounce = gridOf((const void*)I.bounce);
    I.gridShade = I.shade ? gridOf((const void*)I.shade) : 0;
    I.gridShadow = gridOf((const void*)I.shadow);
    I.gridTrace = std::max(gridOf((const void*)I.kn.traceC), gridOf((const void*)I.kn.traceA));
    I.maxGrid = std::max(I.gridBounce, std::max(I.gridShadow, I.gridTrace));
    gridBlocks = I.gridBounce;

    // stack spill beyond the LDS part: the tree bounds the entries a traversal can hold
    I.P.spillCap = stackNeed > I.sl ? stackNeed - I.sl : 0;
    if (I.P.spillCap > 0) {
        uint64_t* sp;
        if (!I.alloc(sp, (size_t)I.maxGrid * kBlock * I.P.spillCap, err)) return false;
        I.P.spill = sp;
    }

    I.N = F.cam.w * F.cam.h;
    I.K = O.batch < 1 ? 1 : O.batch;
    const size_t NK = (size_t)I.N * I.K;
    if (NK > (size_t)0x7fffffff) { err = "width * height * batch exceeds 2^31 paths"; return false; }
    I.NK = (int)NK;
    I.depths = O.bounces + 2;
    if (!I.alloc(I.P.O, NK, err) || !I.alloc(I.P.D, NK, err) || !I.alloc(I.P.T, NK, err) || !I.alloc(I.P.L, NK, err) ||
        !I.alloc(I.P.q[0], NK, err) || !I.alloc(I.P.q[1], NK, err) || !I.alloc(I.P.sO, NK, err) ||
        !I.alloc(I.P.sD, NK, err) || !I.alloc(I.P.sC, NK, err) || !I.alloc(I.P.cnt, (size_t)(I.depths + 1) * kCntStride, err) ||
        !I.alloc(I.accum, (size_t)I.N, err) || !I.alloc(I.tot, 8, err))
        return false;
    if (!O.fuse && !I.alloc(I.P.H, NK, err)) return false;
    HIPCHK(hipHostMalloc((void**)&I.hostTot, 8 * sizeof(unsigned long long), hipHostMallocDefault));
    if (!I.fill(I.accum, (size_t)I.N * 16, err)) return false;
    HIPCHK(hipStreamSynchronize(I.st));
    size_t free1 = 0;
    HIPCHK(hipMemGetInfo(&free1, &total));
    vramMB = (double)(free0 - free1) / 1048576.0;
    return true;
}

bool Renderer::setParams(const FrameParams& F, std::string& err) {
    if (F.cam.w != p_->F.cam.w || F.cam.h != p_->F.cam.h) { err = "setParams cannot change the image size"; return false; }
    if (F.hasN && !p_->S.nrm) { err = "hasN set without normals"; return false; }
    p_->F = F;
    return true;
}

bool Renderer::render(int spp, int sBase, bool clear, FrameStats& S, std::string& err) {
    Impl& I = *p_;
    if (spp <= 0 || spp % I.K) { err = "spp must be a positive multiple of batch"; return false; }
    if (sBase < 0) { err = "sample base must be >= 0"; return false; }
    S = FrameStats();
    I.evUsed = 0;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> bounceEv, shadeEv, shadowEv;
    if (clear && !I.fill(I.accum, (size_t)I.N * 16, err)) return false;
    if (!I.fill(I.tot, 8 * sizeof(unsigned long long), err)) return false;
    HIPCHK(hipEventRecord(I.f0, I.st));
    const int words = (I.depths + 1) * kCntStride;
    for (int s = sBase; s < sBase + spp; s += I.K) {
        k_pass_begin<<<1, 64, 0, I.st>>>(I.P.cnt, words, I.NK);
        for (int depth = 0; depth < I.depths; depth++) {
            hipEvent_t a = nullptr, b = nullptr;
            if (I.O.timeKernels) { a = I.nextEvent(); b = I.nextEvent(); HIPCHK(hipEventRecord(a, I.st)); }
            I.bounce<<<I.gridBounce, kBlock, 0, I.st>>>(I.S, I.P, I.F, depth, s, I.K);
            if (I.shade) {
                if (I.O.timeKernels) { hipEvent_t c = I.nextEvent(); HIPCHK(hipEventRecord(c, I.st)); shadeEv.push_back({ c, nullptr }); }
                I.shade<<<I.gridShade, kBlock, 0, I.st>>>(I.S, I.P, I.F, depth, s, I.K);
            }
            if (I.O.timeKernels) {
                HIPCHK(hipEventRecord(b, I.st)); bounceEv.push_back({ a, b });
                if (I.shade) shadeEv.back().second = b;
            }
            if (I.O.timeKernels) { a = I.nextEvent(); b = I.nextEvent(); HIPCHK(hipEventRecord(a, I.st)); }
            I.shadow<<<I.gridShadow, kBlock, 0, I.st>>>(I.S, I.P, depth);
            if (I.O.timeKernels) { HIPCHK(hipEventRecord(b, I.st)); shadowEv.push_back({ a, b }); }
        }
        k_reduce<<<std::min((I.N + kWave - 1) / kWave, 16384), kWave, 0, I.st>>>(I.P.L, I.accum, I.N, I.K);
        k_pass_end<<<1, 64, 0, I.st>>>(I.P.cnt, I.depths, I.tot);
    }
    HIPCHK(hipGetLastError());
    HIPCHK(hipEventRecord(I.f1, I.st));
    HIPCHK(hipMemcpyAsync(I.hostTot, I.tot, 8 * sizeof(unsigned long long), hipMemcpyDeviceToHost, I.st));
    HIPCHK(hipEventSynchronize(I.f1));
    HIPCHK(hipStreamSynchronize(I.st));
    float ms = 0;
    HIPCHK(hipEventElapsedTime(&ms, I.f0, I.f1));
    S.frameMs = ms;
    for (auto& e : bounceEv) { HIPCHK(hipEventElapsedTime(&ms, e.first, e.second)); S.bounceMs += ms; S.bounceLaunches++; }
    for (auto& e : shadowEv) { HIPCHK(hipEventElapsedTime(&ms, e.first, e.second)); S.shadowMs += ms; S.shadowLaunches++; }
    for (auto& e : shadeEv) { HIPCHK(hipEventElapsedTime(&ms, e.first, e.second)); S.shadeMs += ms; }
    S.traceMs = I.O.timeKernels ? S.bounceMs + S.shadowMs : S.frameMs;
    S.rays = (double)I.hostTot[0]; S.shadowRays = (double)I.hostTot[1];
    S.missRays = (double)I.hostTot[2]; S.unoccluded = (double)I.hostTot[3]; S.contRays = (double)I.hostTot[4];
    S.hitRays = S.rays - S.missRays; S.cameraMiss = (double)I.hostTot[5];
    return true;
}

bool Renderer::readAccum(float* rgba, size_t nPix, std::string& err) {
    if (nPix < (size_t)p_->N) { err = "accumulation buffer too small"; return false; }
    HIPCHK(hipStreamSynchronize(p_->st));
    HIPCHK(hipMemcpy(rgba, p_->accum, (size_t)p_->N * 16, hipMemcpyDeviceToHost));
    return true;
}

bool Renderer::trace(const void* rays, int n, bool anyHit, void* hits, std::string& err) {
    Impl& I = *p_;
    if (n < 0) { err = "negative ray count"; return false; }
    if (n == 0) return true;
    TraceFn fn = anyHit ? I.kn.traceA : I.kn.traceC;
    fn<<<I.gridTrace, kBlock, 0, I.st>>>(I.S, (const Vec4*)rays, n, (Vec4*)hits, I.P.spill, I.P.spillCap);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(I.st));
    return true;
}

bool Renderer::traceHost(const float* rays, int n, bool anyHit, float* hits, std::string& err) {
    Impl& I = *p_;
    if (n < 0) { err = "negative ray count"; return false; }
    if (n == 0) return true;
    if ((size_t)n > I.stageCap) {
        if (I.stageRays) { hipFree(I.stageRays); I.stageRays = nullptr; }
        if (I.stageHits) { hipFree(I.stageHits); I.stageHits = nullptr; }
        I.stageCap = 0;
        HIPCHK(hipMalloc(&I.stageRays, (size_t)n * 32));
        HIPCHK(hipMalloc(&I.stageHits, (size_t)n * 16));
        I.stageCap = (size_t)n;
    }
    HIPCHK(hipMemcpyAsync(I.stageRays, rays, (size_t)n * 32, hipMemcpyHostToDevice, I.st));
    if (!trace(I.stageRays, n, anyHit, I.stageHits, err)) return false;
    HIPCHK(hipMemcpy(hits, I.stageHits, (size_t)n * 16, hipMemcpyDeviceToHost));
    return true;
}

}  // namespace cwp
