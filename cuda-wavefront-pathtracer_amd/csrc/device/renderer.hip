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
#include "device/collapse.h"
#include "device/display.h"
#include "device/hploc.h"
#include "device/kernels.h"
#include "host/bvh.h"
#include "host/envmap.h"
#include "host/scene_check.h"
#include <algorithm>
#include <type_traits>
#include <cstdio>
#include <cstring>

namespace cwp {

using namespace dev;

namespace {

// TLAS input: world boxes of the instances, the 8 corners of the mesh's object box through the forward rows
// (reference: k_inst_boxes, src/render/wavefront.cu:445-451)
__global__ void k_inst_boxes(const InstanceRecord* __restrict__ inst, const Vec4* __restrict__ meshBox, int n, Vec4* lo, Vec4* hi) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const InstanceRecord& I = inst[i];
    const Vec4 bl = meshBox[2 * I.mesh], bh = meshBox[2 * I.mesh + 1];
    Vec3 mn = v3(3.4e38f, 3.4e38f, 3.4e38f), mx = v3(-3.4e38f, -3.4e38f, -3.4e38f);
    for (int c = 0; c < 8; c++) {
        const Vec3 w = xfRowsP(I.fwd, v3(c & 1 ? bh.x : bl.x, c & 2 ? bh.y : bl.y, c & 4 ? bh.z : bl.z));
        mn = v3(w.x < mn.x ? w.x : mn.x, w.y < mn.y ? w.y : mn.y, w.z < mn.z ? w.z : mn.z);
        mx = v3(w.x > mx.x ? w.x : mx.x, w.y > mx.y ? w.y : mx.y, w.z > mx.z ? w.z : mx.z);
    }
    lo[i] = v4(mn, 0.f);
    hi[i] = v4(mx, 0.f);
}

// parents of the H-PLOC nodes (parent[] filled with -1 first), then the deepest leaf's count of internal ancestors:
// the TLAS traversal never holds more stack entries than that
__global__ void k_tlas_parents(const Vec4* __restrict__ hn, int n, int* parent) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n - 1) return;
    const int id = n + j;
    const int a = asInt(hn[2 * id].w), b = asInt(hn[2 * id + 1].w);
    // child ids outside the tree only come from an inconsistent build, which the builder's status reports
    if (a >= 0 && a < 2 * n - 1) parent[a] = id;
    if (b >= 0 && b < 2 * n - 1) parent[b] = id;
}
// 80-byte compressed BVH8 nodes to the device form (deviceBVH8 on the host): one 128-byte line per node, the
// leaf-slot mask (meta bit 7 of each slot) in the top byte of the child base
__global__ void k_bvh8_device(const Vec4* __restrict__ n80, int nn, Vec4* __restrict__ out) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= nn) return;
    const Vec4* n = n80 + (size_t)j * 5;
    Vec4* o = out + (size_t)j * kBvh8Stride;
    const Vec4 f1 = n[1];
    const uint32_t meta[2] = { asUint(f1.z), asUint(f1.w) };
    uint32_t leaf = 0;
    for (int s = 0; s < 8; s++) leaf |= ((meta[s >> 2] >> (8 * (s & 3) + 7)) & 1u) << s;
    o[0] = n[0];
    o[1] = v4(asFloatU(((uint32_t)asInt(f1.x) & 0xFFFFFFu) | (leaf << 24)), f1.y, f1.z, f1.w);
    o[2] = n[2]; o[3] = n[3]; o[4] = n[4];
}

__global__ void k_tlas_depth(const int* __restrict__ parent, int n, int* depth) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int d = 0;
    for (int x = parent[i]; x >= 0 && d < 2 * n; x = parent[x]) d++;   // the bound only guards against a corrupt tree
    atomicMax(depth, d);
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
    FrameParams* dF = nullptr;    // device copy read by the bounce / shade kernels
    // captured frame (graph mode): the launch parameters baked into it and the scene state it was captured with
    hipGraphExec_t graph = nullptr;
    // full integrator on two-level scenes: the frame's emitters must be set with the instances
    bool full = false;
    int numEmInst = 0, numMats = 0, nAlbedo = 0;
    std::vector<Material> hostMats;   // the full integrator's materials as uploaded (emission must not change)
    bool emSet = false;
    int gSpp = 0, gBase = 0, sceneGen = 0, gGen = -1;
    bool gClear = false;
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
    BounceFn bounceTail = nullptr;                 // straggler hand-off: the tail kernels after each bounce and
    ShadowFn shadowTail = nullptr;                 // each shadow launch
    BounceFn bounceTail2 = nullptr;                // second hand-off level (tailLevels 2)
    ShadowFn shadowTail2 = nullptr;
    int gridTail = 0, gridShadowTail = 0;
    FirstFn bounceFirst = nullptr;                 // first-bounce sort fused into depth 0 (single-level scenes)
    int gridFirst = 0;                             // in kFirstTile-thread workgroups
    ShadowFn shadow = nullptr;
    int gridBounce = 0, gridShade = 0, gridShadow = 0, gridTrace = 0, maxGrid = 0, gridSort = 0;
    void* stageRays = nullptr; void* stageHits = nullptr; int* stageInst = nullptr; size_t stageCap = 0;   // traceHost staging
    // two-level scenes: instance records, their world boxes, the TLAS (H-PLOC each frame), BLAS roots per mesh
    bool inst = false, instSet = false;
    int nInst = 0, instCap = 0, nMesh = 0;
    std::vector<int> meshRoot;
    InstanceRecord* dInst = nullptr;
    Vec4 *dLo = nullptr, *dHi = nullptr, *dMeshBox = nullptr, *dTlas = nullptr;
    int *dParent = nullptr, *dDepth = nullptr, *hostDepth = nullptr;
    HplocBuilder tlasB;
    // compressed BVH8 TLAS (BVH8 BLASes): the collapse's 80-byte nodes, their 128-byte device form, leaf instance ids
    bool tlasWide = false;
    CwbvhCollapser tlasC;
    Vec4 *dTlas80 = nullptr, *dTlas8 = nullptr;
    int* dTlasPrim = nullptr;
    hipEvent_t tl0 = nullptr, tl1 = nullptr;

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
    // pass counters: kCntStride per depth (+1), then the continuation sort's bucket counters per depth
    size_t counterWords() const { return (size_t)(depths + 1) * kCntStride + (O.sort ? (size_t)depths * kSortWords : 0); }
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
    if (p_->graph) hipGraphExecDestroy(p_->graph);
    for (void* a : p_->allocs) hipFree(a);
    if (p_->stageRays) hipFree(p_->stageRays);
    if (p_->stageHits) hipFree(p_->stageHits);
    if (p_->stageInst) hipFree(p_->stageInst);
    for (void* a : { (void*)p_->dInst, (void*)p_->dLo, (void*)p_->dHi, (void*)p_->dTlas, (void*)p_->dParent,
                     (void*)p_->dTlas80, (void*)p_->dTlas8, (void*)p_->dTlasPrim })
        if (a) hipFree(a);
    if (p_->hostDepth) hipHostFree(p_->hostDepth);
    if (p_->tl0) hipEventDestroy(p_->tl0);
    if (p_->tl1) hipEventDestroy(p_->tl1);
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
    numCU = prop.multiProcessorCount;
    name = prop.name;
    if (name.empty()) name = prop.gcnArchName;   // some runtimes leave the marketing name empty
    if (name.empty()) name = "gfx (unnamed)";
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
    if (!validMortonMode(O.mortonBits)) { err = "morton_bits must be 0 (adaptive) or 30"; return false; }
    // a frame launches bounces + 2 depths: a deeper path limit would end paths early without an error
    if (F.maxB > O.bounces) { err = "max_bounce (" + std::to_string(F.maxB) + ") exceeds bounces (" + std::to_string(O.bounces) + ")"; return false; }
    I.O = O; I.F = F;
    I.sl = O.stackLds == 8 ? 8 : (O.stackLds == 32 ? 32 : 16);
    I.O.stackLds = I.sl;
    HIPCHK(hipSetDevice(0));
    if (!deviceInfo(numCU, gpuName, err)) return false;
    size_t free0 = 0, total = 0;
    HIPCHK(hipMemGetInfo(&free0, &total));
    HIPCHK(hipStreamCreateWithFlags(&I.st, hipStreamNonBlocking));
    HIPCHK(hipEventCreate(&I.f0)); HIPCHK(hipEventCreate(&I.f1));

    // BVH8 nodes are re-encoded from 80 to 128 bytes: each node then fills exactly one 128-byte cache line instead of
    // straddling two in 3 cases out of 5, and its planes are binary16 (one mixed FMA per plane)
    std::vector<Vec4> devNodes;
    if (S.variant == 8) {
        if (!deviceBVH8(S.nodes, devNodes, err)) return false;
    } else devNodes = S.nodes;
    Vec4 *dn, *dw, *dv, *dnr, *da;
    if (!I.upload(dn, devNodes, err) || !I.upload(dw, S.woop, err) || !I.upload(dv, S.verts, err) ||
        !I.upload(dnr, S.normals, err) || !I.upload(da, S.albedo, err)) return false;
    I.S.nodes = dn; I.S.verts = dv; I.S.nrm = dnr; I.S.albedo = da; I.S.root = S.root;
    {
        const int nt = (int)(S.verts.size() / 3);
        Vec4* dg;
        if (!I.alloc(dg, (size_t)std::max(nt, 1), err)) return false;
        if (nt > 0) k_tri_geo<<<(nt + 255) / 256, 256, 0, I.st>>>(dv, nt, dg);
        HIPCHK(hipGetLastError());
        I.S.geo = dg;
    }
    I.S.tri = O.tri == kTriWoop ? dw : dv;
    if (F.hasN && !dnr) { err = "hasN set without normals"; return false; }
    auto up = [&](auto& p, const auto& v, std::string& e) -> bool {
        using T = typename std::decay_t<decltype(v)>::value_type;
        T* q = nullptr;
        if (!I.upload(q, v, e)) return false;
        p = q;
        return true;
    };
    if (!makeDevLights(S, up, I.S.Lt, err)) return false;

    if (O.fuse != 0 && O.fuse != 1) { err = "fuse must be 0 or 1"; return false; }
    if (O.persistent < 0 || O.persistent > 2) { err = "persistent must be 0, 1 (split kernels) or 2 (fused)"; return false; }
    if (O.persistent == 1 && O.fuse) { err = "persistent traversal 1 runs the split kernels (fuse 0)"; return false; }
    if (O.persistent == 2 && !O.fuse) { err = "persistent traversal 2 runs the fused kernel (fuse 1)"; return false; }
    if (S.variant == 2) pickKernels<2>(O.tri, I.sl, O.bounceWaves, I.kn);
    else if (S.variant == 4) pickKernels<4>(O.tri, I.sl, O.bounceWaves, I.kn);
    else pickKernels<8>(O.tri, I.sl, O.bounceWaves, I.kn);
    I.inst = S.inst != 0;
    if (S.full && (!O.fuse || O.persistent == 1)) {
        err = "the full integrator runs the fused kernel (fuse 1, persistent 0 or 2)"; return false;
    }
    I.full = S.full != 0;
    I.numEmInst = S.inst ? S.numEmInst : 0;
    I.numMats = (int)S.mats.size();
    I.hostMats = S.mats;
    I.nAlbedo = (int)S.albedo.size();
    if (S.full && !I.inst) {
        I.bounce = O.persistent == 2 ? I.kn.bounceRefillFull : I.kn.bounceFull; I.shade = nullptr;
        I.shadow = O.shadowRefill ? I.kn.shadowRefill : I.kn.shadow;
    } else if (I.inst) {
        // two-level scenes run the fused bounce kernel and the one-ray-per-lane shadow kernel
        if (!O.fuse || O.persistent) { err = "two-level scenes need the fused kernel (fuse 1, persistent 0)"; return false; }
        I.bounce = S.full ? I.kn.bounceFullInst : I.kn.bounceInst; I.shade = nullptr; I.shadow = I.kn.shadowInst;
        I.kn.traceC = I.kn.traceCInst; I.kn.traceA = I.kn.traceAInst;
        I.meshRoot = S.meshRoot; I.nMesh = (int)S.meshRoot.size();
        Vec4* mb;
        if (!I.upload(mb, S.meshBox, err)) return false;
        I.dMeshBox = mb;
        HIPCHK(hipEventCreate(&I.tl0)); HIPCHK(hipEventCreate(&I.tl1));
        if (O.tlasWidth != 2 && O.tlasWidth != 8) { err = "TLAS width must be 2 or 8"; return false; }
        I.tlasWide = O.variant == 8 && O.tlasWidth == 8;
        tlasWide = I.tlasWide ? 1 : 0;
        HIPCHK(hipHostMalloc((void**)&I.hostDepth, 2 * sizeof(int), hipHostMallocDefault));   // TLAS depth, build status
        if (!I.alloc(I.dDepth, 1, err)) return false;
    } else {
        I.bounce = O.fuse ? (O.persistent == 2 ? I.kn.bounceRefill : I.kn.bounce) : (O.persistent ? I.kn.tracePaths : I.kn.tracePlain);
        I.shade = O.fuse ? nullptr : k_shade;
        I.shadow = (O.persistent == 1 || O.shadowRefill) ? I.kn.shadowRefill : I.kn.shadow;
    }
    if (O.tail < -1 || O.tail == 1 || O.tail > kWave) { err = "tail must be -1 (auto), 0 (off) or 2..64 lanes"; return false; }
    if (O.tailLevels != 1 && O.tailLevels != 2) { err = "tail levels must be 1 or 2"; return false; }
    // two-level scenes hand off between TLAS node steps, which needs the BVH8 TLAS
    const bool tailable = S.variant == 8 && (!I.inst || I.tlasWide) && O.fuse && !O.persistent;
    // automatic: single-level scenes only (the 1,001-instance frame measured 579 -> 593 ms with it: its hand-off
    // points are TLAS steps, each holding whole BLAS traversals)
    I.O.tail = O.tail == -1 ? (tailable && !I.inst ? kTailDefault : 0) : O.tail;
    if (I.O.tail) {
        if (!tailable) {
            err = "the straggler hand-off (tail) needs a BVH8 scene (two-level: with the BVH8 TLAS) and the fused kernel "
                  "(fuse 1, persistent 0)";
            return false;
        }
        if (I.inst) {
            I.bounce = S.full ? I.kn.bounceHandFullInst : I.kn.bounceHandInst;
            I.bounceTail = S.full ? I.kn.bounceTailFullInst : I.kn.bounceTailInst;
            I.shadow = I.kn.shadowHandInst; I.shadowTail = I.kn.shadowTailInst;
            if (O.tailLevels == 2) {
                I.bounceTail2 = S.full ? I.kn.bounceTail2FullInst : I.kn.bounceTail2Inst;
                I.shadowTail2 = I.kn.shadowTail2Inst;
            }
        } else {
            I.bounce = S.full ? I.kn.bounceHandFull : I.kn.bounceHand;
            I.bounceTail = S.full ? I.kn.bounceTailFull : I.kn.bounceTail;
            if (!O.shadowRefill) { I.shadow = I.kn.shadowHand; I.shadowTail = I.kn.shadowTail; }
            if (O.tailLevels == 2) {
                I.bounceTail2 = S.full ? I.kn.bounceTail2Full : I.kn.bounceTail2;
                if (I.shadowTail) I.shadowTail2 = I.kn.shadowTail2;
            }
        }
    }
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
    if (O.sort == 2 && !I.inst && O.fuse && !O.persistent) {
        I.bounceFirst = S.full ? I.kn.bounceFirstFull : I.kn.bounceFirst;
        int b = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, (const void*)I.bounceFirst, kFirstTile, 0) != hipSuccess || b < 1) b = 1;
        I.gridFirst = std::max(kParts, (b * numCU) / kParts * kParts);
    }
    I.gridTail = I.bounceTail ? gridOf((const void*)I.bounceTail) : 0;
    if (I.bounceTail2) I.gridTail = std::min(I.gridTail, gridOf((const void*)I.bounceTail2));
    I.gridShadowTail = I.shadowTail ? gridOf((const void*)I.shadowTail) : 0;
    I.maxGrid = std::max(std::max(std::max(I.gridBounce, I.gridTail), std::max(I.gridShadow, I.gridShadowTail)),
                         std::max(I.gridTrace, I.gridFirst * (kFirstTile / kBlock)));
    I.gridSort = std::max(kParts, 4 * numCU / kParts * kParts);
    gridBlocks = I.gridBounce;

    // stack spill beyond the LDS part: the tree bounds the entries a traversal can hold (two-level scenes: the
    // deepest accepted TLAS below the deepest BLAS)
    if (I.inst) stackNeed += kTlasMaxDepth;
    I.P.spillCap = stackNeed > I.sl ? stackNeed - I.sl : 0;
    if (I.P.spillCap > 0) {
        uint64_t* sp;
        if (!I.alloc(sp, (size_t)I.maxGrid * kBlock * I.P.spillCap, err)) return false;
        I.P.spill = sp;
    }

    if (!I.alloc(I.dF, 1, err)) return false;
    HIPCHK(hipMemcpy(I.dF, &I.F, sizeof(FrameParams), hipMemcpyHostToDevice));
    I.N = F.cam.w * F.cam.h;
    I.K = O.batch < 1 ? 1 : O.batch;
    const size_t NK = (size_t)I.N * I.K;
    if (NK > (size_t)0x7fffffff) { err = "width * height * batch exceeds 2^31 paths"; return false; }
    I.NK = (int)NK;
    I.depths = O.bounces + 2;
    // ray sets by queue slot: depth d >= 1 reads set d & 1 and appends to the other (set 0 is read from depth 2 on,
    // depth 1 reads set 1)
    if (!I.alloc(I.P.O[0], NK, err) || !I.alloc(I.P.D[0], NK, err) || !I.alloc(I.P.T[0], NK, err) ||
        !I.alloc(I.P.O[1], NK, err) || !I.alloc(I.P.D[1], NK, err) || !I.alloc(I.P.T[1], NK, err) ||
        !I.alloc(I.P.L, NK, err) || !I.alloc(I.P.sO, NK, err) ||
        !I.alloc(I.P.sD, NK, err) || !I.alloc(I.P.sC, NK, err) || !I.alloc(I.P.cnt, I.counterWords(), err) ||
        !I.alloc(I.accum, (size_t)I.N, err) || !I.alloc(I.tot, 8, err))
        return false;
    if (!O.fuse && !I.alloc(I.P.H, NK, err)) return false;
    if (I.O.tail) {   // tail records: slot, hit, node group and the stack (two entries per Vec4), at most lanes - 1 per chunk
        I.P.tailLanes = I.O.tail;
        I.P.tailStride = 3 + (std::max(stackNeed, 1) + 1) / 2;
        const size_t records = ((NK >> 6) + kParts + 1) * (size_t)(I.O.tail - 1);
        if (!I.alloc(I.P.tail, records * (size_t)I.P.tailStride, err)) return false;
        if (I.bounceTail2) {   // the second level's regions follow from the first's record numbering
            const size_t records2 = ((records >> 6) + kParts + 1) * (size_t)(I.O.tail - 1);
            if (!I.alloc(I.P.tail2, records2 * (size_t)I.P.tailStride, err)) return false;
        }
    }
    if (O.sort < 0 || O.sort > 2) { err = "sort must be 0, 1 or 2"; return false; }
    if (O.sort == 2 && (!O.fuse || O.persistent)) { err = "sort 2 (first-bounce sort) needs the fused kernel (fuse 1, persistent 0)"; return false; }
    if (O.sort) {
        if (!I.alloc(I.P.stO, NK, err) || !I.alloc(I.P.stD, NK, err) || !I.alloc(I.P.stT, NK, err) ||
            !I.alloc(I.P.qk, NK, err)) return false;
        I.P.bcnt = I.P.cnt + (size_t)(I.depths + 1) * kCntStride;
    }
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
    if (F.maxB > p_->O.bounces) { err = "max_bounce (" + std::to_string(F.maxB) + ") exceeds bounces (" + std::to_string(p_->O.bounces) + ")"; return false; }
    HIPCHK(hipStreamSynchronize(p_->st));   // queued frames read the previous parameters
    HIPCHK(hipMemcpy(p_->dF, &F, sizeof(FrameParams), hipMemcpyHostToDevice));
    p_->F = F;
    return true;
}

bool Renderer::render(int spp, int sBase, bool clear, FrameStats& S, std::string& err) {
    Impl& I = *p_;
    if (spp <= 0 || spp % I.K) { err = "spp must be a positive multiple of batch"; return false; }
    if (I.inst && !I.instSet) { err = "two-level scene: set the instances of a frame first"; return false; }
    if (I.inst && I.full && I.numEmInst > 0 && !I.emSet) { err = "two-level scene: set the emitters of a frame first"; return false; }
    if (sBase < 0) { err = "sample base must be >= 0"; return false; }
    S = FrameStats();
    I.evUsed = 0;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> bounceEv, shadeEv, shadowEv, sortEv, tailEv;
    const bool timed = I.O.timeKernels && !I.O.graph;
    // the frame's launches: clears (by kernels, never memset nodes), then per pass the counter reset, trace+shade
    // and shadow per bounce, the reduction and the totals
    auto enqueue = [&]() -> bool {
        if (clear && !I.fill(I.accum, (size_t)I.N * 16, err)) return false;
        if (!I.fill(I.tot, 8 * sizeof(unsigned long long), err)) return false;
        const int words = (int)I.counterWords();
        for (int s = sBase; s < sBase + spp; s += I.K) {
            k_pass_begin<<<1, 64, 0, I.st>>>(I.P.cnt, words, I.NK);
            // first-bounce sort: the depth-0 launch stages its continuations, the later ones append directly
            DevPaths P0 = I.P, Pn = I.P;
            if (I.O.sort == 2) { P0.firstSort = I.bounceFirst ? 0 : 1; Pn.qk = nullptr; P0.qk = I.bounceFirst ? nullptr : P0.qk; }
            for (int depth = 0; depth < I.depths; depth++) {
                hipEvent_t a = nullptr, b = nullptr;
                if (timed) { a = I.nextEvent(); b = I.nextEvent(); HIPCHK(hipEventRecord(a, I.st)); }
                const bool first = depth == 0 && I.bounceFirst;   // depth 0 with the first-bounce sort fused in
                if (first) I.bounceFirst<<<I.gridFirst, kFirstTile, 0, I.st>>>(I.S, P0, I.dF, s, I.K);
                else I.bounce<<<I.gridBounce, kBlock, 0, I.st>>>(I.S, depth == 0 ? P0 : Pn, I.dF, depth, s, I.K);
                if (I.bounceTail && !first) {
                    if (timed) { hipEvent_t c = I.nextEvent(); HIPCHK(hipEventRecord(c, I.st)); tailEv.push_back({ c, nullptr }); }
                    I.bounceTail<<<I.gridTail, kBlock, 0, I.st>>>(I.S, depth == 0 ? P0 : Pn, I.dF, depth, s, I.K);
                    if (I.bounceTail2) I.bounceTail2<<<I.gridTail, kBlock, 0, I.st>>>(I.S, depth == 0 ? P0 : Pn, I.dF, depth, s, I.K);
                }
                if (I.shade) {
                    if (timed) { hipEvent_t c = I.nextEvent(); HIPCHK(hipEventRecord(c, I.st)); shadeEv.push_back({ c, nullptr }); }
                    I.shade<<<I.gridShade, kBlock, 0, I.st>>>(I.S, I.P, I.dF, depth, s, I.K);
                }
                if (timed) {
                    HIPCHK(hipEventRecord(b, I.st)); bounceEv.push_back({ a, b });
                    if (I.shade) shadeEv.back().second = b;
                    if (I.bounceTail && !first) tailEv.back().second = b;
                }
                if (timed) { a = I.nextEvent(); b = I.nextEvent(); HIPCHK(hipEventRecord(a, I.st)); }
                I.shadow<<<I.gridShadow, kBlock, 0, I.st>>>(I.S, I.P, depth);
                if (I.shadowTail) {
                    if (timed) { hipEvent_t c = I.nextEvent(); HIPCHK(hipEventRecord(c, I.st)); tailEv.push_back({ c, nullptr }); }
                    I.shadowTail<<<I.gridShadowTail, kBlock, 0, I.st>>>(I.S, I.P, depth);
                    if (I.shadowTail2) I.shadowTail2<<<I.gridShadowTail, kBlock, 0, I.st>>>(I.S, I.P, depth);
                }
                if (timed) {
                    HIPCHK(hipEventRecord(b, I.st)); shadowEv.push_back({ a, b });
                    if (I.shadowTail) tailEv.back().second = b;
                }
                if (I.O.sort == 2 && !I.bounceFirst && depth == 0 && depth + 1 < I.depths) {
                    if (timed) { a = I.nextEvent(); b = I.nextEvent(); HIPCHK(hipEventRecord(a, I.st)); }
                    k_qsort_first<<<I.gridSort, kFirstTile, 0, I.st>>>(P0);
                    if (timed) { HIPCHK(hipEventRecord(b, I.st)); sortEv.push_back({ a, b }); }
                } else if (I.O.sort == 1 && depth + 1 < I.depths) {
                    if (timed) { a = I.nextEvent(); b = I.nextEvent(); HIPCHK(hipEventRecord(a, I.st)); }
                    k_qcount<<<I.gridSort, kSortBlock, 0, I.st>>>(I.P, depth);
                    k_qscatter<<<I.gridSort, kSortBlock, 0, I.st>>>(I.P, depth);
                    if (timed) { HIPCHK(hipEventRecord(b, I.st)); sortEv.push_back({ a, b }); }
                }
            }
            k_reduce<<<std::min((I.N + kWave - 1) / kWave, 16384), kWave, 0, I.st>>>(I.P.L, I.accum, I.N, I.K);
            k_pass_end<<<1, 64, 0, I.st>>>(I.P.cnt, I.depths, I.tot);
        }
        HIPCHK(hipGetLastError());
        return true;
    };
    if (I.O.graph) {
        // one captured graph per (spp, sample base, clear) and scene state; re-captured when any of them changes
        if (!I.graph || I.gSpp != spp || I.gBase != sBase || I.gClear != clear || I.gGen != I.sceneGen) {
            if (I.graph) { HIPCHK(hipGraphExecDestroy(I.graph)); I.graph = nullptr; }
            hipGraph_t g = nullptr;
            HIPCHK(hipStreamBeginCapture(I.st, hipStreamCaptureModeThreadLocal));
            const bool ok = enqueue();
            const hipError_t ec = hipStreamEndCapture(I.st, &g);
            if (!ok) { if (g) hipGraphDestroy(g); return false; }
            HIPCHK(ec);
            const hipError_t ei = hipGraphInstantiate(&I.graph, g, nullptr, nullptr, 0);
            hipGraphDestroy(g);
            if (ei != hipSuccess) { I.graph = nullptr; err = std::string("graph instantiate: ") + hipGetErrorString(ei); return false; }
            I.gSpp = spp; I.gBase = sBase; I.gClear = clear; I.gGen = I.sceneGen;
        }
        HIPCHK(hipEventRecord(I.f0, I.st));
        HIPCHK(hipGraphLaunch(I.graph, I.st));
    } else {
        HIPCHK(hipEventRecord(I.f0, I.st));
        if (!enqueue()) return false;
    }
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
    for (auto& e : sortEv) { HIPCHK(hipEventElapsedTime(&ms, e.first, e.second)); S.sortMs += ms; }
    for (auto& e : tailEv) { HIPCHK(hipEventElapsedTime(&ms, e.first, e.second)); S.tailMs += ms; }
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

const Vec4* Renderer::accumDevice() const { return p_->accum; }

bool Renderer::display(uint32_t* out, int spp, float exposureEV, int mode, std::string& err) {
    if (!displayLaunch(p_->accum, out, p_->N, spp, exposureEV, mode, (void*)p_->st, err)) return false;
    HIPCHK(hipStreamSynchronize(p_->st));
    return true;
}

bool Renderer::trace(const void* rays, int n, bool anyHit, void* hits, std::string& err, int* inst) {
    Impl& I = *p_;
    if (n < 0) { err = "negative ray count"; return false; }
    if (n == 0) return true;
    if (I.inst && !I.instSet) { err = "two-level scene: set the instances of a frame first"; return false; }
    TraceFn fn = anyHit ? I.kn.traceA : I.kn.traceC;
    fn<<<I.gridTrace, kBlock, 0, I.st>>>(I.S, (const Vec4*)rays, n, (Vec4*)hits, inst, I.P.spill, I.P.spillCap);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(I.st));
    return true;
}

bool Renderer::traceHost(const float* rays, int n, bool anyHit, float* hits, std::string& err, int* inst) {
    Impl& I = *p_;
    if (n < 0) { err = "negative ray count"; return false; }
    if (n == 0) return true;
    if ((size_t)n > I.stageCap) {
        if (I.stageRays) { hipFree(I.stageRays); I.stageRays = nullptr; }
        if (I.stageHits) { hipFree(I.stageHits); I.stageHits = nullptr; }
        if (I.stageInst) { hipFree(I.stageInst); I.stageInst = nullptr; }
        I.stageCap = 0;
        HIPCHK(hipMalloc(&I.stageRays, (size_t)n * 32));
        HIPCHK(hipMalloc(&I.stageHits, (size_t)n * 16));
        HIPCHK(hipMalloc((void**)&I.stageInst, (size_t)n * 4));
        I.stageCap = (size_t)n;
    }
    HIPCHK(hipMemcpyAsync(I.stageRays, rays, (size_t)n * 32, hipMemcpyHostToDevice, I.st));
    if (!trace(I.stageRays, n, anyHit, I.stageHits, err, inst ? I.stageInst : nullptr)) return false;
    HIPCHK(hipMemcpy(hits, I.stageHits, (size_t)n * 16, hipMemcpyDeviceToHost));
    if (inst) HIPCHK(hipMemcpy(inst, I.stageInst, (size_t)n * 4, hipMemcpyDeviceToHost));
    return true;
}

bool Renderer::setMaterials(const Material* m, int n, std::string& err) {
    Impl& I = *p_;
    if (!m || n < 1) { err = "setMaterials: no materials"; return false; }
    if (n != (I.full ? I.numMats : I.nAlbedo)) {
        err = "setMaterials: the scene has " + std::to_string(I.full ? I.numMats : I.nAlbedo) + " materials"; return false;
    }
    for (int k = 0; k < n; k++) {
        const Material& x = m[k];
        if (!(x.ior > 0.f) || !(x.metallic >= 0.f && x.metallic <= 1.f) || !(x.transmission >= 0.f && x.transmission <= 1.f) ||
            !(x.roughness >= 0.f && x.roughness <= 1.f)) { err = "setMaterials: material " + std::to_string(k) + " out of range"; return false; }
        if (I.full && memcmp(&x.emission, &I.hostMats[k].emission, sizeof(Vec3)) != 0) {
            err = "setMaterials: material " + std::to_string(k) + " changes its emission (the light tables depend on it)"; return false;
        }
    }
    std::vector<Vec4> alb(n);
    for (int k = 0; k < n; k++) alb[k] = v4(m[k].baseColor, 0.f);
    HIPCHK(hipStreamSynchronize(I.st));   // queued frames read the previous materials
    HIPCHK(hipMemcpy((void*)I.S.albedo, alb.data(), (size_t)n * sizeof(Vec4), hipMemcpyHostToDevice));
    if (I.full) {
        HIPCHK(hipMemcpy((void*)I.S.Lt.mats, m, (size_t)n * sizeof(Material), hipMemcpyHostToDevice));
        I.hostMats.assign(m, m + n);
    }
    return true;
}

bool Renderer::setEnvironment(const Vec4* texels, size_t count, float envRotDeg, bool sunOn, std::string& err) {
    Impl& I = *p_;
    DevLights& L = I.S.Lt;
    if (!I.full) { err = "setEnvironment: the scene renders with the legacy integrator"; return false; }
    EnvMap E;
    if (texels) {
        if (!L.hasEnv) { err = "setEnvironment: the scene has no environment map"; return false; }
        if (count != (size_t)L.envW * L.envH) {
            err = "setEnvironment: the environment has " + std::to_string((size_t)L.envW * L.envH) + " texels"; return false;
        }
        std::vector<float> rgb(count * 3);
        for (size_t t = 0; t < count; t++) { rgb[3 * t] = texels[t].x; rgb[3 * t + 1] = texels[t].y; rgb[3 * t + 2] = texels[t].z; }
        buildEnvMap(rgb, L.envW, L.envH, 1.f, E);   // texels and the sampling distribution built from them
    }
    HIPCHK(hipStreamSynchronize(I.st));   // queued frames read the previous environment
    if (texels) {
        HIPCHK(hipMemcpy((void*)L.envTex, E.texels.data(), E.texels.size() * sizeof(Vec4), hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy((void*)L.envMarg, E.marg.data(), E.marg.size() * sizeof(float), hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy((void*)L.envCond, E.cond.data(), E.cond.size() * sizeof(float), hipMemcpyHostToDevice));
    }
    L.envRot = envRotDeg * 3.14159265f / 180.f;
    // light-type selection: equal shares of the light types present (the rule of the scene's creation)
    const bool tri = L.nEm > 0 && L.lightMode != kLightBSDF, env = L.hasEnv && L.lightMode != kLightBSDF;
    const int nt = (sunOn ? 1 : 0) + (tri ? 1 : 0) + (env ? 1 : 0);
    L.pSun = sunOn && nt ? 1.f / (float)nt : 0.f; L.pTri = tri ? 1.f / (float)nt : 0.f; L.pEnv = env ? 1.f / (float)nt : 0.f;
    I.sceneGen++;   // captured frames hold the scene record by value
    return true;
}

bool Renderer::setEmitters(const Vec4* emTris, int nEm, const float* emCdf, const float* matLightPdf, int nMats,
                           std::string& err) {
    Impl& I = *p_;
    if (!I.inst || !I.full) { err = "setEmitters: the scene is not a two-level full-integrator scene"; return false; }
    if (nEm != I.numEmInst || nMats != I.numMats) { err = "setEmitters: the scene has " + std::to_string(I.numEmInst) +
        " instanced emitters and " + std::to_string(I.numMats) + " materials"; return false; }
    if (nEm == 0) { I.emSet = true; return true; }
    if (!emTris || !emCdf || !matLightPdf) { err = "setEmitters: null array"; return false; }
    // the kernels trust the material ids and the CDF
    if (emCdf[0] != 0.f || emCdf[nEm] != 1.f) { err = "setEmitters: the CDF must run from 0 to 1"; return false; }
    for (int e = 0; e < nEm; e++) {
        const int m = asInt(emTris[3 * e + 1].w);
        if (m < 0 || m >= nMats) { err = "setEmitters: emitter " + std::to_string(e) + " has a material id out of range"; return false; }
        if (emCdf[e + 1] < emCdf[e]) { err = "setEmitters: the CDF decreases"; return false; }
    }
    HIPCHK(hipStreamSynchronize(I.st));   // queued frames read the previous frame's emitters
    HIPCHK(hipMemcpy((void*)I.S.Lt.emTris, emTris, (size_t)nEm * 3 * sizeof(Vec4), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy((void*)I.S.Lt.emCdf, emCdf, ((size_t)nEm + 1) * sizeof(float), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy((void*)I.S.Lt.matLightPdf, matLightPdf, (size_t)nMats * sizeof(float), hipMemcpyHostToDevice));
    I.emSet = true;
    return true;
}

bool Renderer::setInstances(const InstanceRecord* rec, int n, double& tlasMs, std::string& err) {
    Impl& I = *p_;
    if (!I.inst) { err = "setInstances: the scene is not two-level"; return false; }
    if (n < 1 || !rec) { err = "setInstances: no instances"; return false; }
    if (I.tlasWide && n > (1 << 24)) { err = "setInstances: more than 2^24 instances for a BVH8 TLAS"; return false; }
    // the kernels trust the records: every instance must name a mesh and that mesh's BLAS root
    for (int i = 0; i < n; i++)
        if (rec[i].mesh < 0 || rec[i].mesh >= I.nMesh || rec[i].root != I.meshRoot[rec[i].mesh]) {
            err = "setInstances: instance " + std::to_string(i) + " has a bad mesh or BLAS root"; return false;
        }
    HIPCHK(hipStreamSynchronize(I.st));   // the previous frame may still read the TLAS
    // from here on the previous instance set is gone: a failure below leaves the scene without instances (render then
    // reports an error), and frames captured with the old TLAS are never replayed
    I.instSet = false;
    I.sceneGen++;
    if (n > I.instCap) {
        for (void* a : { (void*)I.dInst, (void*)I.dLo, (void*)I.dHi, (void*)I.dTlas, (void*)I.dParent, (void*)I.dTlas80,
                         (void*)I.dTlas8, (void*)I.dTlasPrim }) if (a) hipFree(a);
        I.dInst = nullptr; I.dLo = I.dHi = I.dTlas = I.dTlas80 = I.dTlas8 = nullptr; I.dParent = I.dTlasPrim = nullptr;
        I.instCap = 0;
        HIPCHK(hipMalloc((void**)&I.dInst, (size_t)n * sizeof(InstanceRecord)));
        HIPCHK(hipMalloc((void**)&I.dLo, (size_t)n * 16));
        HIPCHK(hipMalloc((void**)&I.dHi, (size_t)n * 16));
        HIPCHK(hipMalloc((void**)&I.dTlas, (size_t)std::max(n - 1, 1) * 64));
        HIPCHK(hipMalloc((void**)&I.dParent, (size_t)(2 * n - 1) * 4));
        if (I.tlasWide) {   // a wide tree over n primitives has at most n nodes
            HIPCHK(hipMalloc((void**)&I.dTlas80, (size_t)n * 5 * sizeof(Vec4)));
            HIPCHK(hipMalloc((void**)&I.dTlas8, (size_t)n * kBvh8Stride * sizeof(Vec4)));
            HIPCHK(hipMalloc((void**)&I.dTlasPrim, (size_t)n * sizeof(int)));
        }
        if (!I.tlasB.reserve(n, err)) return false;
        I.instCap = n;
    }
    HIPCHK(hipMemcpyAsync(I.dInst, rec, (size_t)n * sizeof(InstanceRecord), hipMemcpyHostToDevice, I.st));
    HIPCHK(hipEventRecord(I.tl0, I.st));
    k_inst_boxes<<<(n + 255) / 256, 256, 0, I.st>>>(I.dInst, I.dMeshBox, n, I.dLo, I.dHi);
    int hroot = 0, wideNodes = 0, wideDepth = 0;
    if (!I.tlasB.build(I.dLo, I.dHi, n, I.st, hroot, err, I.O.mortonBits)) return false;
    if (I.tlasWide) {
        // the collapse checks the H-PLOC tree itself and synchronises with the stream between its levels
        if (!I.tlasC.run(I.tlasB.nodes(), n, I.dTlas80, I.dTlasPrim, I.st, wideNodes, wideDepth, err)) return false;
        k_bvh8_device<<<(wideNodes + 255) / 256, 256, 0, I.st>>>(I.dTlas80, wideNodes, I.dTlas8);
    } else if (!hplocToBVH2(I.tlasB.nodes(), n, I.dTlas, I.st, err)) return false;
    HIPCHK(hipEventRecord(I.tl1, I.st));
    if (!I.fill(I.dDepth, 4, err)) return false;
    if (!I.tlasWide) {
        k_fill<<<std::max(1, std::min((2 * n - 1 + 255) / 256, 4096)), 256, 0, I.st>>>((uint32_t*)I.dParent, (size_t)(2 * n - 1), 0xffffffffu);
        if (n > 1) k_tlas_parents<<<(n - 1 + 255) / 256, 256, 0, I.st>>>(I.tlasB.nodes(), n, I.dParent);
        k_tlas_depth<<<(n + 255) / 256, 256, 0, I.st>>>(I.dParent, n, I.dDepth);
    }
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(I.hostDepth, I.dDepth, 4, hipMemcpyDeviceToHost, I.st));
    HIPCHK(hipMemcpyAsync(I.hostDepth + 1, I.tlasB.status(), 4, hipMemcpyDeviceToHost, I.st));
    HIPCHK(hipStreamSynchronize(I.st));
    if (I.hostDepth[1] != 0) {
        err = "setInstances: inconsistent TLAS build (status " + std::to_string(I.hostDepth[1]) + ")";
        return false;
    }
    float ms = 0;
    HIPCHK(hipEventElapsedTime(&ms, I.tl0, I.tl1));
    tlasMs = ms;
    this->tlasMs = ms;
    if (I.tlasWide) *I.hostDepth = wideDepth;   // one stacked group per wide level at most
    if (*I.hostDepth + 1 > kTlasMaxDepth) {
        I.instSet = false;
        err = "setInstances: TLAS depth " + std::to_string(*I.hostDepth) + " exceeds the traversal stack";
        return false;
    }
    if (I.tlasWide) { I.S.tlas = I.dTlas8; I.S.tlasRoot = 0; I.S.tlasWide = 1; I.S.tlasPrim = I.dTlasPrim; }
    else { I.S.tlas = I.dTlas; I.S.tlasRoot = hplocBVH2Root(n); I.S.tlasWide = 0; I.S.tlasPrim = nullptr; }
    I.S.inst = I.dInst;
    I.sceneGen++;   // captured frames hold the scene record by value
    I.nInst = n; I.instSet = true;
    return true;
}

}  // namespace cwp
