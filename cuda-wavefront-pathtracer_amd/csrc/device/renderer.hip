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
#include "device/kernels.h"
#include "host/scene_check.h"
#include <algorithm>
#include <cstdio>

namespace cwp {

using namespace dev;

namespace {

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
