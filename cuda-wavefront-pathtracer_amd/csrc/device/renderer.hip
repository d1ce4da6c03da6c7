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
#include <hip/hip_runtime.h>
#include <algorithm>
#include <cstdio>

namespace cwp {

namespace {

constexpr int kWave = 64;
constexpr int kBlock = 64;       // one wave per workgroup: LDS stacks are sized per wave
constexpr int kParts = 8;        // queue partitions, one per XCD (workgroup w runs on XCD w % 8)
// Per-depth counters (kCntStride ints): [x] path-queue length of partition x, [kParts + x] shadow rays produced by
// partition x, [2 kParts] misses, [2 kParts + 1] unoccluded shadow rays.
constexpr int kCntStride = 32;
constexpr int kCntMiss = 2 * kParts, kCntUnocc = 2 * kParts + 1;

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
    int* q[2];            // ping-pong queues of path ids: depth d reads q[d & 1]
    Vec4 *sO, *sD, *sC;   // shadow rays of the current bounce: (origin, path id bits), (dir, tmax), (contribution)
    int* cnt;             // kCntStride counters per depth
    void* spill;          // traversal-stack spill slab, spillCap entries per thread
    int spillCap;
};

template <int V> struct StackEntry;
template <> struct StackEntry<2> { using type = int; };
template <> struct StackEntry<8> { using type = Group; };

template <int V, bool ANY, int TRI, class St>
__device__ __forceinline__ bool traceScene(const DevScene& S, const RayPre& r, Hit& h, St& st) {
    if constexpr (V == 2) return traceBVH2<ANY, TRI>(S.nodes, S.tri, S.root, r, h, st);
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

__device__ __forceinline__ void waveCount(uint64_t mask, int* counter) {
    if (mask && laneId() == __builtin_ctzll(mask)) atomicAdd(counter, (int)__builtin_popcountll(mask));
}

// start of partition `part` in a region laid out as partitions of the given lengths
__device__ __forceinline__ int partStart(const int* len, int part) {
    int s = 0;
    for (int y = 0; y < part; y++) s += len[y];
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
// One bounce of the legacy integrator over the queue of depth `depth` (depth 0: every path id, camera rays made here).
// The queue is split into kParts partitions, one per XCD; the waves of XCD x walk partition x, and append their
// continuations to partition x of the next queue and their shadow rays to partition x of the shadow queue, so a
// path stays on one XCD (and in its L2) for its whole life. Partition x of the depth-(d+1) queue starts where the
// depth-d partitions before x end, which bounds it by the rays partition x traced.
// WAVES: minimum waves per SIMD the register allocation must allow (1 = unconstrained).
template <int V, int TRI, int SL, int WAVES>
__global__ void __launch_bounds__(kBlock, WAVES) k_bounce(DevScene S, DevPaths P, FrameParams F, int depth, int smpBase, int K) {
    using E = typename StackEntry<V>::type;
    __shared__ E lds[SL > 0 ? SL * kBlock : 1];
    const int lane = laneId();
    const int gtid = blockIdx.x * kBlock + threadIdx.x;
    LaneStack<E, SL> st{ lds + threadIdx.x, kBlock, (E*)P.spill + (size_t)gtid * P.spillCap, 0 };
    const int part = (int)(blockIdx.x % kParts), lw = (int)(blockIdx.x / kParts), nw = (int)(gridDim.x / kParts);
    int* cnt = P.cnt + depth * kCntStride;
    const int n = cnt[part];
    const int inStart = depth == 0 ? partStart(cnt, part) : partStart(cnt - kCntStride, part);
    const int outStart = partStart(cnt, part);
    const int* qIn = P.q[depth & 1];
    int* qOut = P.q[(depth + 1) & 1];
    for (int base = lw * kWave; base < n; base += nw * kWave) {
        const int k = base + lane;
        const bool act = k < n;
        bool cont = false, shadow = false, miss = false;
        int i = 0;
        Vec3 po = v3(0.f, 0.f, 0.f), nd = po, Tn = po, sdir = po, scon = po;
        if (act) {
            i = depth == 0 ? inStart + k : qIn[inStart + k];
            Vec3 o, d;
            float tmin;
            if (depth == 0) { cameraRay(F, i / K, (uint32_t)(smpBase + i % K), o, d); tmin = F.nearT; }
            else { o = xyz(P.O[i]); d = xyz(P.D[i]); tmin = 0.f; }
            const RayPre r = prepareRay<TRI>(o, d, tmin);
            Hit h; h.t = 1e30f; h.u = 0.f; h.v = 0.f; h.tri = -1;
            st.sp = 0;
            traceScene<V, false, TRI>(S, r, h, st);
            // throughput, pixel and sample are fetched / recomputed after the traversal so that no registers hold
            // them across it
            const Vec3 T = depth == 0 ? v3(1.f, 1.f, 1.f) : xyz(P.T[i]);
            const int pix = i / K;
            const uint32_t smp = (uint32_t)(smpBase + i % K);
            o = r.o; d = r.d;
            if (h.tri < 0) {
                miss = true;
                Vec3 c = T * F.sky;
                if (depth == 0) P.L[i] = v4(c, 0.f);
                else { Vec4 l = P.L[i]; l.x += c.x; l.y += c.y; l.z += c.z; P.L[i] = l; }
            } else {
                if (depth == 0) P.L[i] = v4(0.f, 0.f, 0.f, 0.f);
                if (depth <= F.maxB) {
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
                    po = p + ng * F.eps;
                    const uint32_t dd = (uint32_t)depth + 1u;
                    const Vec3 l = uniformCone(F.sunTo, F.sunCos, rnd((uint32_t)pix, smp, dd, 2, F.seed), rnd((uint32_t)pix, smp, dd, 3, F.seed));
                    const float cl = dot(ns, l);
                    if (cl > 0.f && dot(ng, l) > 0.f) { shadow = true; sdir = l; scon = T * alb * F.sunE * (cl * kInvPi); }
                    nd = cosineHemisphere(ns, rnd((uint32_t)pix, smp, dd, 0, F.seed), rnd((uint32_t)pix, smp, dd, 1, F.seed));
                    if (dot(nd, ng) > 0.f) { cont = true; Tn = T * alb; }
                }
            }
        }
        const uint64_t mc = __ballot(cont), ms = __ballot(shadow), mm = __ballot(miss);
        if (mc) {
            const int pos = outStart + waveAppend(mc, cnt + kCntStride + part);
            if (cont) { qOut[pos] = i; P.O[i] = v4(po, 0.f); P.D[i] = v4(nd, 0.f); P.T[i] = v4(Tn, 0.f); }
        }
        if (ms) {
            const int pos = outStart + waveAppend(ms, cnt + kParts + part);
            if (shadow) { P.sO[pos] = v4(po, asFloat(i)); P.sD[pos] = v4(sdir, 1e30f); P.sC[pos] = v4(scon, 0.f); }
        }
        waveCount(mm, cnt + kCntMiss);
    }
}

// Shadow rays of bounce `depth` (partition per XCD as in k_bounce): any hit; unoccluded contributions are added to
// the path's radiance. The any-hit traversal fits 64 registers without spilling: 8 waves per SIMD.
template <int V, int TRI, int SL>
__global__ void __launch_bounds__(kBlock, 8) k_shadow(DevScene S, DevPaths P, int depth) {
    using E = typename StackEntry<V>::type;
    __shared__ E lds[SL > 0 ? SL * kBlock : 1];
    const int gtid = blockIdx.x * kBlock + threadIdx.x;
    LaneStack<E, SL> st{ lds + threadIdx.x, kBlock, (E*)P.spill + (size_t)gtid * P.spillCap, 0 };
    const int part = (int)(blockIdx.x % kParts), lw = (int)(blockIdx.x / kParts), nw = (int)(gridDim.x / kParts);
    int* cnt = P.cnt + depth * kCntStride;
    const int n = cnt[kParts + part];
    const int start = partStart(cnt, part);
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
        waveCount(__ballot(vis), cnt + kCntUnocc);
    }
}

// Closest or any hit of explicit rays (parity and kernel-level entry point).
template <int V, int TRI, bool ANY, int SL>
__global__ void __launch_bounds__(kBlock, 8) k_trace(DevScene S, const Vec4* __restrict__ rays, int n, Vec4* __restrict__ hits, void* spill, int spillCap) {
    using E = typename StackEntry<V>::type;
    __shared__ E lds[SL > 0 ? SL * kBlock : 1];
    const int gtid = blockIdx.x * kBlock + threadIdx.x;
    LaneStack<E, SL> st{ lds + threadIdx.x, kBlock, (E*)spill + (size_t)gtid * spillCap, 0 };
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
        cnt[x] = (int)((x + 1) * NK / kParts - x * NK / kParts);
    }
}

// per-pass totals: [0] closest rays, [1] shadow rays, [2] misses, [3] unoccluded, [4] continuations,
// [5] camera-ray misses
__global__ void k_pass_end(const int* cnt, int depths, unsigned long long* tot) {
    if (threadIdx.x != 0) return;
    unsigned long long a = 0, b = 0, m = 0, u = 0, c = 0, m0 = (unsigned)cnt[kCntMiss];
    for (int d = 0; d < depths; d++) {
        const int* cd = cnt + d * kCntStride;
        unsigned long long q = 0;
        for (int x = 0; x < kParts; x++) { q += (unsigned)cd[x]; b += (unsigned)cd[kParts + x]; }
        a += q; if (d > 0) c += q;
        m += (unsigned)cd[kCntMiss]; u += (unsigned)cd[kCntUnocc];
    }
    tot[0] += a; tot[1] += b; tot[2] += m; tot[3] += u; tot[4] += c; tot[5] += m0;
}

// accum[pix] += sum over the K samples of the pass (path ids pix*K .. pix*K + K-1), in sample order
__global__ void k_reduce(const Vec4* __restrict__ L, Vec4* __restrict__ accum, int N, int K) {
    for (int pix = blockIdx.x * blockDim.x + threadIdx.x; pix < N; pix += gridDim.x * blockDim.x) {
        float r = 0.f, g = 0.f, b = 0.f;
        const Vec4* l = L + (size_t)pix * K;
        for (int s = 0; s < K; s++) { const Vec4 v = l[s]; r += v.x; g += v.y; b += v.z; }
        Vec4 a = accum[pix]; a.x += r; a.y += g; a.z += b; accum[pix] = a;
    }
}

__global__ void k_fill(uint32_t* p, size_t words, uint32_t v) {
    for (size_t j = blockIdx.x * (size_t)blockDim.x + threadIdx.x; j < words; j += (size_t)gridDim.x * blockDim.x) p[j] = v;
}

// ------------------------------------------------------------------------------------------------------- dispatch
using BounceFn = void (*)(DevScene, DevPaths, FrameParams, int, int, int);
using ShadowFn = void (*)(DevScene, DevPaths, int);
using TraceFn = void (*)(DevScene, const Vec4*, int, Vec4*, void*, int);

template <int V, int TRI, int SL>
BounceFn pickBounce(int waves) {
    if (waves >= 8) return k_bounce<V, TRI, SL, 8>;
    if (waves >= 6) return k_bounce<V, TRI, SL, 6>;
    return k_bounce<V, TRI, SL, 1>;
}
template <int V, int TRI>
void pickBySL(int sl, int waves, BounceFn& b, ShadowFn& s, TraceFn& tc, TraceFn& ta) {
    switch (sl) {
        case 8: b = pickBounce<V, TRI, 8>(waves); s = k_shadow<V, TRI, 8>; tc = k_trace<V, TRI, false, 8>; ta = k_trace<V, TRI, true, 8>; break;
        case 32: b = pickBounce<V, TRI, 32>(waves); s = k_shadow<V, TRI, 32>; tc = k_trace<V, TRI, false, 32>; ta = k_trace<V, TRI, true, 32>; break;
        default: b = pickBounce<V, TRI, 16>(waves); s = k_shadow<V, TRI, 16>; tc = k_trace<V, TRI, false, 16>; ta = k_trace<V, TRI, true, 16>; break;
    }
}
template <int V>
void pickByTri(int tri, int sl, int waves, BounceFn& b, ShadowFn& s, TraceFn& tc, TraceFn& ta) {
    if (tri == kTriMT) pickBySL<V, kTriMT>(sl, waves, b, s, tc, ta);
    else if (tri == kTriWatertight) pickBySL<V, kTriWatertight>(sl, waves, b, s, tc, ta);
    else pickBySL<V, kTriWoop>(sl, waves, b, s, tc, ta);
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
    BounceFn bounce = nullptr; ShadowFn shadow = nullptr; TraceFn traceC = nullptr, traceA = nullptr;
    int gridBounce = 0, gridShadow = 0, gridTrace = 0, maxGrid = 0;
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

bool Renderer::init(const SceneUpload& S, const FrameParams& F, const RenderOptions& O, std::string& err) {
    Impl& I = *p_;
    if (S.variant != 2 && S.variant != 8) { err = "variant must be 2 or 8"; return false; }
    if (O.tri < kTriMT || O.tri > kTriWatertight) { err = "unknown triangle test"; return false; }
    if (O.tri == kTriWoop && S.woop.size() != S.verts.size()) { err = "Woop rows missing"; return false; }
    if (F.cam.w <= 0 || F.cam.h <= 0) { err = "image size must be positive"; return false; }
    if (S.verts.empty() || S.nodes.empty()) { err = "empty scene"; return false; }
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

    if (S.variant == 2) pickByTri<2>(O.tri, I.sl, O.bounceWaves, I.bounce, I.shadow, I.traceC, I.traceA);
    else pickByTri<8>(O.tri, I.sl, O.bounceWaves, I.bounce, I.shadow, I.traceC, I.traceA);
    auto gridOf = [&](const void* fn) {
        int b = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, fn, kBlock, 0) != hipSuccess || b < 1) b = 1;
        if (O.blocksPerCU > 0) b = std::min(b, O.blocksPerCU);
        return std::max(kParts, (b * numCU) / kParts * kParts);   // a whole number of waves per partition
    };
    I.gridBounce = gridOf((const void*)I.bounce);
    I.gridShadow = gridOf((const void*)I.shadow);
    I.gridTrace = std::max(gridOf((const void*)I.traceC), gridOf((const void*)I.traceA));
    I.maxGrid = std::max(I.gridBounce, std::max(I.gridShadow, I.gridTrace));
    gridBlocks = I.gridBounce;

    // stack spill beyond the LDS part: the tree bounds the entries a traversal can hold
    I.P.spillCap = S.stackNeed > I.sl ? S.stackNeed - I.sl : 0;
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
    std::vector<std::pair<hipEvent_t, hipEvent_t>> bounceEv, shadowEv;
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
            if (I.O.timeKernels) { HIPCHK(hipEventRecord(b, I.st)); bounceEv.push_back({ a, b }); }
            if (I.O.timeKernels) { a = I.nextEvent(); b = I.nextEvent(); HIPCHK(hipEventRecord(a, I.st)); }
            I.shadow<<<I.gridShadow, kBlock, 0, I.st>>>(I.S, I.P, depth);
            if (I.O.timeKernels) { HIPCHK(hipEventRecord(b, I.st)); shadowEv.push_back({ a, b }); }
        }
        k_reduce<<<std::min((I.N + 255) / 256, 8192), 256, 0, I.st>>>(I.P.L, I.accum, I.N, I.K);
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
    TraceFn fn = anyHit ? I.traceA : I.traceC;
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
