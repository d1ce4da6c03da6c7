// device/batch.hip - BatchRenderer: resident scenes and one wavefront pass over many jobs (see batch.h).
#include "device/batch.h"
#include "device/kernels.h"
#include "host/bvh.h"
#include "host/scene_check.h"
#include <algorithm>
#include <cstddef>
#include <type_traits>

namespace cwp {

using namespace dev;

namespace {

#define HIPCHK(x)                                                                                                      \
    do {                                                                                                               \
        hipError_t e_ = (x);                                                                                           \
        if (e_ != hipSuccess) { err = std::string(#x) + ": " + hipGetErrorString(e_); return false; }                  \
    } while (0)

struct SceneDev {
    bool alive = false;
    std::vector<void*> bufs;
    DevScene S{};
    size_t bytes = 0;
    int stackNeed = 0;
};

constexpr int kTableSlots = 4;   // job tables in flight (passes may be queued without host synchronisation)
constexpr int kRollSlots = 8;    // step tables in flight of a rolling batch (even: a slot always holds steps of one parity)

}  // namespace

struct BatchRenderer::Impl {
    BatchOptions O;
    hipStream_t st = nullptr;
    hipStream_t cst = nullptr;   // image read-back of rolling batches (does not wait for queued steps)
    hipEvent_t e0 = nullptr, e1 = nullptr;
    std::vector<SceneDev> scenes;
    DevPaths P{};
    std::vector<void*> allocs;
    Kernels kn;
    int numCU = 0, grid = 0, gridFull = 0, redGrid = 0, depths = 0, sl = 8, spillCap = 0;
    unsigned long long *tot = nullptr, *hostTot = nullptr;
    DevJobTable* dTable[kTableSlots] = {};
    DevJobTable* hTable[kTableSlots] = {};
    hipEvent_t tableDone[kTableSlots] = {};
    bool tableUsed[kTableSlots] = {};
    int slot = 0;
    // one graph per table slot, valid while the path state it was captured with is (pathGen)
    hipGraphExec_t graph[kTableSlots] = {};
    int graphGen[kTableSlots] = {};
    int pathGen = 1;
    // rolling batches: step tables, their completion events and graphs (one per slot)
    int gridRoll = 0, gridRollFull = 0;
    DevStepTable* dStep[kRollSlots] = {};
    DevStepTable* hStep[kRollSlots] = {};
    hipEvent_t stepDone[kRollSlots] = {};
    hipGraphExec_t rollGraph[kRollSlots] = {};
    int rollGraphGen[kRollSlots] = {};
    long long steps = 0;   // steps queued since rollBegin

    template <class T> bool alloc(T*& p, size_t count, std::string& err) {
        void* q = nullptr;
        HIPCHK(hipMalloc(&q, std::max<size_t>(count * sizeof(T), 16)));
        allocs.push_back(q); p = (T*)q;
        return true;
    }
    bool fill(void* p, size_t bytes, uint32_t v, std::string& err) {
        const size_t words = bytes / 4;
        const int g = (int)std::max<size_t>(1, std::min<size_t>((words + 255) / 256, 4096));
        k_fill<<<g, 256, 0, st>>>((uint32_t*)p, words, v);
        HIPCHK(hipGetLastError());
        return true;
    }
    // the per-lane spill area must hold the deepest resident scene's stack beyond the LDS entries
    bool ensureSpill(int stackNeed, std::string& err) {
        const int cap = stackNeed > sl ? stackNeed - sl : 0;
        if (cap <= spillCap) return true;
        HIPCHK(hipStreamSynchronize(st));
        if (P.spill) { hipFree(P.spill); P.spill = nullptr; }
        HIPCHK(hipMalloc(&P.spill, (size_t)std::max(std::max(grid, gridFull), std::max(gridRoll, gridRollFull)) * kBlock * cap *
                                   sizeof(uint64_t)));
        spillCap = cap; P.spillCap = cap;
        pathGen++;   // captured graphs hold the old spill and tail pointers
        return ensureTail(err);
    }
    // hand-off records (slot, hit, node group, then the stack two entries per Vec4) for the deepest stack the lanes can
    // hold, sl + spillCap entries; at most lanes - 1 records per 64-entry chunk of a partition's input
    bool ensureTail(std::string& err) {
        if (!O.tail) return true;
        const int stride = 3 + (sl + spillCap + 1) / 2;
        if (P.tail && stride <= P.tailStride) return true;
        HIPCHK(hipStreamSynchronize(st));
        if (P.tail) { hipFree(P.tail); P.tail = nullptr; }
        if (P.tail2) { hipFree(P.tail2); P.tail2 = nullptr; }
        const size_t records = ((O.maxPaths >> 6) + kParts + 1) * (size_t)(O.tail - 1);
        HIPCHK(hipMalloc(&P.tail, records * (size_t)stride * sizeof(Vec4)));
        if (O.tailLevels == 2) {   // the rolling steps' first-tail records are the second level's input (at most
                                   // lanes - 1 records per 64 of them)
            const size_t records2 = ((records >> 6) + kParts + 1) * (size_t)(O.tail - 1);
            HIPCHK(hipMalloc(&P.tail2, records2 * (size_t)stride * sizeof(Vec4)));
        }
        P.tailStride = stride; P.tailLanes = O.tail;
        pathGen++;
        return true;
    }
    // The shadow launches hand off too with the full integrator only. Same-box measurements: the full-integrator
    // serving sweep 2.72 -> 2.84 Grays/s with it at 500 jobs/s (shadow rays to emitters and the environment), the legacy
    // job stream 5.52 -> 5.48 (sun shadow rays: the shadow tails cost 100.5 ms against 88.8 ms of plain shadow launches).
    bool shadowTail() const { return O.tail && O.full; }
    // the launches of a pass reading table slot s
    void enqueue(int s) {
        k_pass_begin_jobs<<<1, 64, 0, st>>>(P.cnt, (depths + 1) * kCntStride, dTable[s]);
        const int gb = O.full ? gridFull : grid;
        for (int depth = 0; depth < depths; depth++) {
            if (O.tail) {   // the hand-off kernel, then its tail (same grid: within the spill area's lanes)
                (O.full ? kn.bounceJobsHandFull : kn.bounceJobsHand)<<<gb, kBlock, 0, st>>>(dTable[s], P, depth);
                (O.full ? kn.bounceJobsTailFull : kn.bounceJobsTail)<<<gb, kBlock, 0, st>>>(dTable[s], P, depth);
            } else {
                (O.full ? kn.bounceJobsFull : kn.bounceJobs)<<<gb, kBlock, 0, st>>>(dTable[s], P, depth);
            }
            if (shadowTail()) {
                kn.shadowJobsHand<<<grid, kBlock, 0, st>>>(dTable[s], P, depth);
                kn.shadowJobsTail<<<grid, kBlock, 0, st>>>(dTable[s], P, depth);
            } else {
                kn.shadowJobs<<<grid, kBlock, 0, st>>>(dTable[s], P, depth);
            }
        }
        k_reduce_jobs<<<redGrid, kWave, 0, st>>>(dTable[s], P.L);
        k_pass_end<<<1, 64, 0, st>>>(P.cnt, depths, tot);
    }
    // the launches of a rolling step reading step-table slot s (step parity = s & 1)
    void enqueueStep(int s, int step) {
        k_step_begin<<<1, 64, 0, st>>>(dStep[s], P.cnt, step);
        const int gb = O.full ? gridRollFull : gridRoll;
        if (O.tail) {
            kn.bounceRollHand[O.full ? 1 : 0][O.compact]<<<gb, kBlock, 0, st>>>(dStep[s], P, step);
            kn.bounceRollTail[O.full ? 1 : 0][O.compact]<<<gb, kBlock, 0, st>>>(dStep[s], P, step);
            if (P.tail2) kn.bounceRollTail2[O.full ? 1 : 0][O.compact]<<<gb, kBlock, 0, st>>>(dStep[s], P, step);
        } else {
            const RollFn b = O.compact ? (O.full ? kn.bounceRollFullLane : kn.bounceRollLane) : (O.full ? kn.bounceRollFull : kn.bounceRoll);
            b<<<gb, kBlock, 0, st>>>(dStep[s], P, step);
        }
        if (shadowTail()) {
            kn.shadowRollHand<<<gridRoll, kBlock, 0, st>>>(dStep[s], P, step);
            kn.shadowRollTail<<<gridRoll, kBlock, 0, st>>>(dStep[s], P, step);
        } else {
            kn.shadowRoll<<<gridRoll, kBlock, 0, st>>>(dStep[s], P, step);
        }
        k_reduce_roll<<<redGrid, kWave, 0, st>>>(dStep[s], P.L);
        k_step_end<<<1, 64, 0, st>>>(P.cnt, step, tot);
    }
    bool captureStep(int s, std::string& err) {
        if (rollGraph[s] && rollGraphGen[s] == pathGen) return true;
        if (rollGraph[s]) { HIPCHK(hipGraphExecDestroy(rollGraph[s])); rollGraph[s] = nullptr; }
        hipGraph_t g = nullptr;
        HIPCHK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
        enqueueStep(s, s);
        const hipError_t e = hipGetLastError();
        HIPCHK(hipStreamEndCapture(st, &g));
        if (e != hipSuccess) { hipGraphDestroy(g); err = std::string("batch: step capture: ") + hipGetErrorString(e); return false; }
        const hipError_t ei = hipGraphInstantiate(&rollGraph[s], g, nullptr, nullptr, 0);
        hipGraphDestroy(g);
        if (ei != hipSuccess) { rollGraph[s] = nullptr; err = std::string("batch: graph instantiate: ") + hipGetErrorString(ei); return false; }
        rollGraphGen[s] = pathGen;
        return true;
    }
    bool capture(int s, std::string& err) {
        if (graph[s] && graphGen[s] == pathGen) return true;
        if (graph[s]) { HIPCHK(hipGraphExecDestroy(graph[s])); graph[s] = nullptr; }
        hipGraph_t g = nullptr;
        HIPCHK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
        enqueue(s);
        const hipError_t e = hipGetLastError();
        HIPCHK(hipStreamEndCapture(st, &g));
        if (e != hipSuccess) { hipGraphDestroy(g); err = std::string("batch: pass capture: ") + hipGetErrorString(e); return false; }
        const hipError_t ei = hipGraphInstantiate(&graph[s], g, nullptr, nullptr, 0);
        hipGraphDestroy(g);
        if (ei != hipSuccess) { graph[s] = nullptr; err = std::string("batch: graph instantiate: ") + hipGetErrorString(ei); return false; }
        graphGen[s] = pathGen;
        return true;
    }
};

BatchRenderer::BatchRenderer() : p_(new Impl) {}
BatchRenderer::~BatchRenderer() {
    if (!p_) return;
    if (p_->st) hipStreamSynchronize(p_->st);
    for (SceneDev& s : p_->scenes) for (void* b : s.bufs) hipFree(b);
    for (void* a : p_->allocs) hipFree(a);
    if (p_->P.spill) hipFree(p_->P.spill);
    if (p_->P.tail) hipFree(p_->P.tail);
    if (p_->P.tail2) hipFree(p_->P.tail2);
    for (int k = 0; k < kTableSlots; k++) {
        if (p_->graph[k]) hipGraphExecDestroy(p_->graph[k]);
        if (p_->hTable[k]) hipHostFree(p_->hTable[k]);
        if (p_->tableDone[k]) hipEventDestroy(p_->tableDone[k]);
    }
    for (int k = 0; k < kRollSlots; k++) {
        if (p_->rollGraph[k]) hipGraphExecDestroy(p_->rollGraph[k]);
        if (p_->hStep[k]) hipHostFree(p_->hStep[k]);
        if (p_->stepDone[k]) hipEventDestroy(p_->stepDone[k]);
    }
    if (p_->hostTot) hipHostFree(p_->hostTot);
    if (p_->e0) hipEventDestroy(p_->e0);
    if (p_->e1) hipEventDestroy(p_->e1);
    if (p_->cst) { hipStreamSynchronize(p_->cst); hipStreamDestroy(p_->cst); }
    if (p_->st) hipStreamDestroy(p_->st);
    delete p_;
}

const BatchOptions& BatchRenderer::options() const { return p_->O; }

bool BatchRenderer::init(const BatchOptions& O, std::string& err) {
    Impl& I = *p_;
    if (!nodeFloat4s(O.variant)) { err = "batch: variant must be 2, 4 or 8"; return false; }
    if (O.tri < kTriMT || O.tri > kTriWatertight) { err = "batch: unknown triangle test"; return false; }
    if (O.stackLds != 8 && O.stackLds != 16 && O.stackLds != 32) { err = "batch: stack_lds must be 8, 16 or 32"; return false; }
    if (O.maxJobs < 1 || O.maxJobs > kMaxJobs) { err = "batch: max jobs must be in [1, " + std::to_string(kMaxJobs) + "]"; return false; }
    if (O.maxPaths < 1 || O.maxPaths > (size_t)0x7fffffff) { err = "batch: max paths must be in [1, 2^31)"; return false; }
    if (O.bounces < 0 || O.bounces > 64) { err = "batch: bounces must be in [0, 64]"; return false; }
    if (O.raysPerLane != 1 && O.raysPerLane != 2 && O.raysPerLane != 4) { err = "batch: rays per lane must be 1, 2 or 4"; return false; }
    if (O.compact != 0 && O.compact != 1) { err = "batch: compaction must be 0 (per wave) or 1 (per lane)"; return false; }
    if (O.tail < -1 || O.tail == 1 || O.tail > kWave) { err = "batch: tail must be -1 (auto), 0 (off) or 2..64 lanes"; return false; }
    if (O.tail > 0 && O.variant != 8) { err = "batch: the straggler hand-off (tail) resumes BVH8 traversals only"; return false; }
    if (O.tailLevels != 1 && O.tailLevels != 2) { err = "batch: tail levels must be 1 or 2"; return false; }
    I.O = O; I.sl = O.stackLds;
    I.O.tail = O.tail == -1 ? (O.variant == 8 ? kTailDefault : 0) : O.tail;
    std::string name;
    if (!deviceInfo(I.numCU, name, err)) return false;
    HIPCHK(hipSetDevice(0));
    size_t free0 = 0, total = 0;
    HIPCHK(hipMemGetInfo(&free0, &total));
    HIPCHK(hipStreamCreateWithFlags(&I.st, hipStreamNonBlocking));
    HIPCHK(hipStreamCreateWithFlags(&I.cst, hipStreamNonBlocking));
    HIPCHK(hipEventCreate(&I.e0)); HIPCHK(hipEventCreate(&I.e1));
    if (O.variant == 2) pickKernels<2>(O.tri, I.sl, 8, I.kn);
    else if (O.variant == 4) pickKernels<4>(O.tri, I.sl, 8, I.kn);
    else pickKernels<8>(O.tri, I.sl, 8, I.kn);
    int b = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, (const void*)I.kn.bounceJobs, kBlock, 0) != hipSuccess || b < 1) b = 1;
    I.grid = std::max(kParts, (b * I.numCU / O.raysPerLane) / kParts * kParts);
    int bf = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&bf, (const void*)I.kn.bounceJobsFull, kBlock, 0) != hipSuccess || bf < 1) bf = 1;
    I.gridFull = std::max(kParts, (bf * I.numCU / O.raysPerLane) / kParts * kParts);
    int br = 0, brf = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&br, (const void*)I.kn.bounceRoll, kBlock, 0) != hipSuccess || br < 1) br = 1;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&brf, (const void*)I.kn.bounceRollFull, kBlock, 0) != hipSuccess || brf < 1) brf = 1;
    I.gridRoll = std::max(kParts, (br * I.numCU / O.raysPerLane) / kParts * kParts);
    I.gridRollFull = std::max(kParts, (brf * I.numCU / O.raysPerLane) / kParts * kParts);
    gridBlocks = I.grid;
    I.redGrid = (int)std::min<size_t>(16384, O.maxPaths / kWave + (size_t)O.maxJobs);
    I.depths = O.bounces + 2;
    const size_t n = O.maxPaths;
    if (!I.alloc(I.P.O[0], n, err) || !I.alloc(I.P.D[0], n, err) || !I.alloc(I.P.T[0], n, err) ||
        !I.alloc(I.P.O[1], n, err) || !I.alloc(I.P.D[1], n, err) || !I.alloc(I.P.T[1], n, err) || !I.alloc(I.P.L, n, err) ||
        !I.alloc(I.P.sO, n, err) || !I.alloc(I.P.sD, n, err) ||
        !I.alloc(I.P.sC, n, err) || !I.alloc(I.P.cnt, (size_t)(I.depths + 1) * kCntStride, err) || !I.alloc(I.tot, 8, err))
        return false;
    HIPCHK(hipHostMalloc((void**)&I.hostTot, 8 * sizeof(unsigned long long), hipHostMallocDefault));
    if (!I.fill(I.tot, 8 * sizeof(unsigned long long), 0u, err) || !I.ensureTail(err)) return false;
    for (int k = 0; k < kTableSlots; k++) {
        if (!I.alloc(I.dTable[k], 1, err)) return false;
        HIPCHK(hipHostMalloc((void**)&I.hTable[k], sizeof(DevJobTable), hipHostMallocDefault));
        HIPCHK(hipEventCreate(&I.tableDone[k]));
    }
    for (int k = 0; k < kRollSlots; k++) {
        if (!I.alloc(I.dStep[k], 1, err)) return false;
        HIPCHK(hipHostMalloc((void**)&I.hStep[k], sizeof(DevStepTable), hipHostMallocDefault));
        HIPCHK(hipEventCreateWithFlags(&I.stepDone[k], hipEventDisableTiming));
    }
    HIPCHK(hipStreamSynchronize(I.st));
    size_t free1 = 0;
    HIPCHK(hipMemGetInfo(&free1, &total));
    vramMB = (double)(free0 - free1) / 1048576.0;
    return true;
}

int BatchRenderer::addScene(const SceneUpload& S, std::string& err) {
    Impl& I = *p_;
    int need = 0;
    if (!checkScene(S, need, err)) return -1;
    if (S.variant != I.O.variant) { err = "batch: the scene's BVH layout differs from the batch's"; return -1; }
    if (S.inst) { err = "batch: two-level scenes are not batched"; return -1; }
    if ((S.full != 0) != (I.O.full != 0)) { err = "batch: the scene's integrator differs from the batch's"; return -1; }
    if (I.O.tri == kTriWoop && S.woop.size() != S.verts.size()) { err = "batch: Woop rows missing"; return -1; }
    SceneDev D;
    auto up = [&](const std::vector<Vec4>& v, const Vec4*& dst) {
        if (v.empty()) { dst = nullptr; return true; }
        void* q = nullptr;
        if (hipMalloc(&q, v.size() * sizeof(Vec4)) != hipSuccess) { err = "batch: out of device memory"; return false; }
        D.bufs.push_back(q); D.bytes += v.size() * sizeof(Vec4);
        if (hipMemcpy(q, v.data(), v.size() * sizeof(Vec4), hipMemcpyHostToDevice) != hipSuccess) { err = "batch: upload failed"; return false; }
        dst = (const Vec4*)q;
        return true;
    };
    std::vector<Vec4> nodes;
    if (S.variant == 8) {   // the 128-byte device encoding, as the renderer's
        if (!deviceBVH8(S.nodes, nodes, err)) return -1;
    } else nodes = S.nodes;
    const Vec4 *dn, *dw, *dv, *dnr, *da;
    static const std::vector<Vec4> none;
    const bool ok = up(nodes, dn) && up(I.O.tri == kTriWoop ? S.woop : none, dw) && up(S.verts, dv) && up(S.normals, dnr) &&
                    up(S.albedo, da);
    auto upAny = [&](auto& p, const auto& v, std::string& e) -> bool {
        if (v.empty()) { p = nullptr; return true; }
        using T = typename std::decay_t<decltype(v)>::value_type;
        void* q = nullptr;
        if (hipMalloc(&q, v.size() * sizeof(T)) != hipSuccess) { e = "batch: out of device memory"; return false; }
        D.bufs.push_back(q); D.bytes += v.size() * sizeof(T);
        if (hipMemcpy(q, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice) != hipSuccess) { e = "batch: upload failed"; return false; }
        p = (T*)q;
        return true;
    };
    if (!ok || !makeDevLights(S, upAny, D.S.Lt, err)) { for (void* b : D.bufs) hipFree(b); return -1; }
    D.S.nodes = dn; D.S.verts = dv; D.S.nrm = dnr; D.S.albedo = da; D.S.root = S.root;
    {   // per-triangle shading record (cross(e1, e2), material id), computed on the device from the corners
        const int nt = (int)(S.verts.size() / 3);
        void* q = nullptr;
        if (hipMalloc(&q, (size_t)std::max(nt, 1) * sizeof(Vec4)) != hipSuccess) {
            err = "batch: out of device memory"; for (void* b : D.bufs) hipFree(b); return -1;
        }
        D.bufs.push_back(q); D.bytes += (size_t)std::max(nt, 1) * sizeof(Vec4);
        if (nt > 0) k_tri_geo<<<(nt + 255) / 256, 256, 0, I.st>>>(dv, nt, (Vec4*)q);
        if (hipGetLastError() != hipSuccess || hipStreamSynchronize(I.st) != hipSuccess) {
            err = "batch: shading-record kernel failed"; for (void* b : D.bufs) hipFree(b); return -1;
        }
        D.S.geo = (const Vec4*)q;
    }
    D.S.tri = I.O.tri == kTriWoop ? dw : dv;
    D.S.tlas = nullptr; D.S.tlasRoot = 0; D.S.inst = nullptr;
    D.stackNeed = need; D.alive = true;
    if (!I.ensureSpill(need, err)) { for (void* b : D.bufs) hipFree(b); return -1; }
    for (size_t k = 0; k < I.scenes.size(); k++)
        if (!I.scenes[k].alive) { I.scenes[k] = D; return (int)k; }
    I.scenes.push_back(D);
    return (int)I.scenes.size() - 1;
}

void BatchRenderer::removeScene(int id) {
    Impl& I = *p_;
    if (id < 0 || id >= (int)I.scenes.size() || !I.scenes[id].alive) return;
    hipStreamSynchronize(I.st);   // queued passes may still read it
    for (void* b : I.scenes[id].bufs) hipFree(b);
    I.scenes[id] = SceneDev();
}

size_t BatchRenderer::uploadBytes(const SceneUpload& S, int tri) {
    const size_t nodes = S.variant == 8 ? S.nodes.size() / 5 * kBvh8Stride : S.nodes.size();
    size_t b = (nodes + (tri == kTriWoop ? S.woop.size() : 0) + S.verts.size() + S.verts.size() / 3 + S.normals.size() +
                S.albedo.size()) * sizeof(Vec4);
    if (S.full) {
        b += S.mats.size() * sizeof(Material) + (S.triLightPdf.size() + S.emTri.size() + S.emCdf.size()) * 4;
        if (S.env) b += S.env->texels.size() * sizeof(Vec4) + (S.env->marg.size() + S.env->cond.size()) * 4;
    }
    return b;
}

size_t BatchRenderer::sceneBytes(int id) const {
    const Impl& I = *p_;
    return id >= 0 && id < (int)I.scenes.size() && I.scenes[id].alive ? I.scenes[id].bytes : 0;
}

bool BatchRenderer::allocImage(int nPix, void** img, std::string& err) {
    *img = nullptr;
    if (nPix < 1) { err = "batch: empty image"; return false; }
    HIPCHK(hipMalloc(img, (size_t)nPix * 16));
    return clearImage(*img, nPix, err);
}
void BatchRenderer::freeImage(void* img) {
    if (!img) return;
    hipStreamSynchronize(p_->st);
    hipFree(img);
}
bool BatchRenderer::clearImage(void* img, int nPix, std::string& err) { return p_->fill(img, (size_t)nPix * 16, 0u, err); }
bool BatchRenderer::readImage(const void* img, int nPix, float* rgba, std::string& err) {
    HIPCHK(hipMemcpyAsync(rgba, img, (size_t)nPix * 16, hipMemcpyDeviceToHost, p_->st));
    HIPCHK(hipStreamSynchronize(p_->st));
    return true;
}
bool BatchRenderer::synchronize(std::string& err) { HIPCHK(hipStreamSynchronize(p_->st)); return true; }

bool BatchRenderer::pass(const std::vector<BatchJob>& jobs, bool sync, FrameStats& st, std::string& err) {
    Impl& I = *p_;
    st = FrameStats();
    const int nj = (int)jobs.size();
    if (nj < 1) return true;
    if (nj > I.O.maxJobs) { err = "batch: too many jobs in one pass"; return false; }
    // the job table of this pass, in a slot whose previous upload has completed
    const int s = I.slot;
    I.slot = (I.slot + 1) % kTableSlots;
    if (I.tableUsed[s]) HIPCHK(hipEventSynchronize(I.tableDone[s]));
    DevJobTable* T = I.hTable[s];
    long long paths = 0;
    int groups = 0;
    for (int j = 0; j < nj; j++) {
        const BatchJob& B = jobs[j];
        if (B.scene < 0 || B.scene >= (int)I.scenes.size() || !I.scenes[B.scene].alive) { err = "batch: job " + std::to_string(j) + " names no resident scene"; return false; }
        const SceneDev& D = I.scenes[B.scene];
        const int N = B.F.cam.w * B.F.cam.h;
        if (N < 1 || B.K < 1 || !B.accum || B.smpBase < 0) { err = "batch: job " + std::to_string(j) + " has no pixels, samples or image"; return false; }
        if (B.F.hasN && !D.S.nrm) { err = "batch: job " + std::to_string(j) + " shades with normals its scene lacks"; return false; }
        if (B.F.maxB > I.O.bounces) { err = "batch: job " + std::to_string(j) + " bounces deeper than the batch allows"; return false; }
        DevJob& J = T->jobs[j];
        J.S = D.S; J.F = B.F; J.accum = (Vec4*)B.accum;
        J.pathBase = (int)paths; J.N = N; J.K = B.K; J.smpBase = B.smpBase; J.grpBase = groups;
        paths += (long long)N * B.K;
        groups += (N + kWave - 1) / kWave;
        if (paths > (long long)I.O.maxPaths) { err = "batch: the pass exceeds the path-state slots"; return false; }
    }
    T->nJobs = nj; T->groups = groups; T->paths = paths;
    if (I.O.useGraph && !I.capture(s, err)) return false;
    HIPCHK(hipMemcpyAsync(I.dTable[s], T, offsetof(DevJobTable, jobs) + (size_t)nj * sizeof(DevJob), hipMemcpyHostToDevice, I.st));
    HIPCHK(hipEventRecord(I.tableDone[s], I.st));
    I.tableUsed[s] = true;
    HIPCHK(hipEventRecord(I.e0, I.st));
    if (I.O.useGraph) HIPCHK(hipGraphLaunch(I.graph[s], I.st));
    else I.enqueue(s);
    HIPCHK(hipEventRecord(I.e1, I.st));
    HIPCHK(hipGetLastError());
    if (!sync) return true;
    HIPCHK(hipStreamSynchronize(I.st));
    float ms = 0;
    HIPCHK(hipEventElapsedTime(&ms, I.e0, I.e1));
    st.frameMs = ms; st.traceMs = ms;
    return true;
}

bool BatchRenderer::rollBegin(std::string& err) {
    Impl& I = *p_;
    // both counter blocks of the step ring start empty (no continuations before the first step)
    if (!I.fill(I.P.cnt, (size_t)2 * kCntStride * sizeof(int), 0u, err)) return false;
    I.steps = 0;
    return true;
}

bool BatchRenderer::rollStep(const std::vector<RollJob>& live, const std::vector<char>& isNew,
                             const std::vector<char>& isDone, long long& stepId, std::string& err) {
    Impl& I = *p_;
    const int nj = (int)live.size();
    if (nj > I.O.maxJobs) { err = "batch: too many live jobs in one step"; return false; }
    if (isNew.size() != live.size() || isDone.size() != live.size()) { err = "batch: step flags do not match the jobs"; return false; }
    const long long step = I.steps;
    const int s = (int)(step % kRollSlots);
    if (step >= kRollSlots) HIPCHK(hipEventSynchronize(I.stepDone[s]));   // the slot's previous step is done
    DevStepTable* T = I.hStep[s];
    // jobs by path base (the kernels find a path's job by binary search); new and done lists in that order
    std::vector<int> order(nj);
    for (int k = 0; k < nj; k++) order[k] = k;
    std::sort(order.begin(), order.end(), [&](int a, int b) { return live[a].pathBase < live[b].pathBase; });
    long long cam = 0;
    int groups = 0, nNew = 0, nDone = 0;
    long long lastEnd = 0;
    for (int r = 0; r < nj; r++) {
        const RollJob& B = live[order[r]];
        if (B.scene < 0 || B.scene >= (int)I.scenes.size() || !I.scenes[B.scene].alive) { err = "batch: a live job names no resident scene"; return false; }
        const SceneDev& D = I.scenes[B.scene];
        const int N = B.F.cam.w * B.F.cam.h;
        if (N < 1 || B.K < 1 || !B.accum || B.smpBase < 0 || B.pathBase < 0) { err = "batch: a live job has no pixels, samples, image or paths"; return false; }
        if (B.pathBase < lastEnd || (size_t)B.pathBase + (size_t)N * B.K > I.O.maxPaths) { err = "batch: live jobs' path ranges overlap or exceed the slots"; return false; }
        if (B.F.hasN && !D.S.nrm) { err = "batch: a live job shades with normals its scene lacks"; return false; }
        if (B.F.maxB > I.O.bounces) { err = "batch: a live job bounces deeper than the batch allows"; return false; }
        lastEnd = (long long)B.pathBase + (long long)N * B.K;
        DevJob& J = T->jobs[r];
        J.S = D.S; J.F = B.F; J.accum = (Vec4*)B.accum;
        J.pathBase = B.pathBase; J.N = N; J.K = B.K; J.smpBase = B.smpBase; J.grpBase = 0;
        if (isNew[order[r]]) { T->newJob[nNew] = r; T->camFirst[nNew] = (int)cam; nNew++; cam += (long long)N * B.K; }
        if (isDone[order[r]]) { T->doneJob[nDone] = r; T->doneGrp[nDone] = groups; nDone++; groups += (N + kWave - 1) / kWave; }
    }
    T->nJobs = nj; T->nNew = nNew; T->nDone = nDone; T->doneGroups = groups; T->camPaths = cam;
    if (I.O.useGraph && !I.captureStep(s, err)) return false;
    HIPCHK(hipMemcpyAsync(I.dStep[s], T, offsetof(DevStepTable, jobs) + (size_t)nj * sizeof(DevJob), hipMemcpyHostToDevice, I.st));
    if (I.O.useGraph) HIPCHK(hipGraphLaunch(I.rollGraph[s], I.st));
    else I.enqueueStep(s, s);
    HIPCHK(hipEventRecord(I.stepDone[s], I.st));
    HIPCHK(hipGetLastError());
    stepId = step;
    I.steps++;
    return true;
}

bool BatchRenderer::stepFinished(long long stepId) const {
    const Impl& I = *p_;
    if (stepId < I.steps - kRollSlots) return true;   // its slot was reused, which waited for it
    return hipEventQuery(I.stepDone[stepId % kRollSlots]) == hipSuccess;
}

bool BatchRenderer::stepSlotFree() const {
    const Impl& I = *p_;
    return I.steps < kRollSlots || hipEventQuery(I.stepDone[I.steps % kRollSlots]) == hipSuccess;
}

bool BatchRenderer::fetchImage(const void* img, int nPix, float* rgba, std::string& err) {
    Impl& I = *p_;
    HIPCHK(hipMemcpyAsync(rgba, img, (size_t)nPix * 16, hipMemcpyDeviceToHost, I.cst));
    HIPCHK(hipStreamSynchronize(I.cst));
    return true;
}

bool BatchRenderer::clearImageAsync(void* img, int nPix, std::string& err) { return p_->fill(img, (size_t)nPix * 16, 0u, err); }

bool BatchRenderer::counters(FrameStats& t, bool reset, std::string& err) {
    Impl& I = *p_;
    t = FrameStats();
    HIPCHK(hipMemcpyAsync(I.hostTot, I.tot, 8 * sizeof(unsigned long long), hipMemcpyDeviceToHost, I.st));
    HIPCHK(hipStreamSynchronize(I.st));
    t.rays = (double)I.hostTot[0]; t.shadowRays = (double)I.hostTot[1];
    t.missRays = (double)I.hostTot[2]; t.unoccluded = (double)I.hostTot[3]; t.contRays = (double)I.hostTot[4];
    t.hitRays = t.rays - t.missRays; t.cameraMiss = (double)I.hostTot[5];
    if (reset && !I.fill(I.tot, 8 * sizeof(unsigned long long), 0u, err)) return false;
    return true;
}

}  // namespace cwp
