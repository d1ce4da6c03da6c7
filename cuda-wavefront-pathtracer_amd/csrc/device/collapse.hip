// device/collapse.hip - H-PLOC nodes -> compressed BVH8 on the device (see collapse.h).
#include "device/collapse.h"
#include <hip/hip_runtime.h>
#include <algorithm>
#include <cmath>

namespace cwp {

namespace {

constexpr int kPMax = 3;                 // primitives per wide leaf (the collapse's pMax)
constexpr float kCNode = 1.f, kCPrim = 0.3f;
constexpr unsigned char kDecLeaf = 0, kDecNode = 1, kDecPrev = 2, kDecDist = 10;
constexpr float kInf = 3e38f;
constexpr int kWaitVmcnt0 = 0x0F70;      // s_waitcnt vmcnt(0)

constexpr int kLevelBatch = 8;   // wide levels launched between host checks

// scratch arrays carved out of one allocation (binary nodes: N = 2n - 1; wide nodes: at most n)
struct Scratch {
    int* parent; int* pcount; int* flag; int* pcnt;     // per binary node: parent, parent count, climb arrivals, primitives
    float* cost;                                        // per binary node, 8 floats: cost as 1..7 wide roots (index 0 unused)
    unsigned char* dec;                                 // per binary node, 8 bytes: [0] split of 8 roots, [1..7] decisions
    int* wBin; int* wSlot; int* wChild; int* wNInt; int* wLp; int* wT; int* wTP; int* wS; int* wP; int* wIdx;
    unsigned char* wLeaf;
    int* ctl;                                           // [0] error, [1] wide node count, [2] max leaf primitives
    int* lv;                                            // level starts: level l is wide nodes [lv[l], lv[l + 1])
};

__device__ __forceinline__ Vec3 boxLo(const Vec4* hn, int h) { return xyz(hn[2 * h]); }
__device__ __forceinline__ Vec3 boxHi(const Vec4* hn, int h) { return xyz(hn[2 * h + 1]); }
__device__ __forceinline__ int leftOf(const Vec4* hn, int h) { return asInt(hn[2 * h].w); }
__device__ __forceinline__ int rightOf(const Vec4* hn, int h) { return asInt(hn[2 * h + 1].w); }
// AABB::area of the host: 0 for an empty box, else 2 (dx dy + dy dz + dz dx)
__device__ __forceinline__ float areaOf(Vec3 lo, Vec3 hi) {
    if (hi.x < lo.x) return 0.f;
    const Vec3 d = hi - lo;
    return 2.f * (d.x * d.y + d.y * d.z + d.z * d.x);
}

// --------------------------------------------------------------------------------------------------- checks
__global__ void k_cb_check(const Vec4* __restrict__ hn, int n, Scratch W) {
    const int N = 2 * n - 1;
    for (int h = blockIdx.x * blockDim.x + threadIdx.x; h < N; h += gridDim.x * blockDim.x) {
        const int l = leftOf(hn, h), r = rightOf(hn, h);
        if (h < n) {
            if (l < 0 || l >= n || r != -1) atomicOr(&W.ctl[0], 1);
        } else if (l < 0 || r < 0 || l >= h || r >= h) {
            atomicOr(&W.ctl[0], 2);
        } else {
            W.parent[l] = h; W.parent[r] = h;
            atomicAdd(&W.pcount[l], 1); atomicAdd(&W.pcount[r], 1);
        }
    }
}
__global__ void k_cb_check2(int n, Scratch W) {
    const int N = 2 * n - 1;
    for (int h = blockIdx.x * blockDim.x + threadIdx.x; h < N; h += gridDim.x * blockDim.x)
        if (W.pcount[h] != (h == N - 1 ? 0 : 1)) atomicOr(&W.ctl[0], 4);
}

// --------------------------------------------------------------------------------------------------- cost tables
// Costs and decisions cross waves through device-coherent accesses (relaxed device-scope atomics), ordered by a
// vmcnt(0) wait before each arrival counter, as in the H-PLOC climb.
__device__ __forceinline__ void stF(float* p, float v) { __hip_atomic_store((uint32_t*)p, __float_as_uint(v), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); }
__device__ __forceinline__ float ldF(const float* p) { return __uint_as_float(__hip_atomic_load((const uint32_t*)p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)); }
__device__ __forceinline__ void stI(int* p, int v) { __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); }
__device__ __forceinline__ int ldI(const int* p) { return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); }
__device__ __forceinline__ void stD(unsigned char* p, const unsigned char d[8]) {
    const uint64_t v = (uint64_t)d[0] | ((uint64_t)d[1] << 8) | ((uint64_t)d[2] << 16) | ((uint64_t)d[3] << 24) |
                       ((uint64_t)d[4] << 32) | ((uint64_t)d[5] << 40) | ((uint64_t)d[6] << 48) | ((uint64_t)d[7] << 56);
    __hip_atomic_store((uint64_t*)p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// One lane per leaf climbs to the root; the second lane to arrive at a node computes its table from both children
// (collapseCWBVH's bottom-up pass: leaves cost A count cPrim at every width; an internal node's cost as i roots is
// the best split of i roots between its children, or the cost as fewer roots; as 1 root it is the better of a leaf
// (at most pMax primitives) and an internal wide node over the best split of 8 roots).
__global__ void __launch_bounds__(256) k_cb_dp(const Vec4* __restrict__ hn, int n, Scratch W) {
    const int h0 = blockIdx.x * blockDim.x + threadIdx.x;
    if (h0 >= n) return;
    const int root = 2 * n - 2;
    {
        const float cl = areaOf(boxLo(hn, h0), boxHi(hn, h0)) * (float)1 * kCPrim;
        unsigned char d[8] = { 1, kDecLeaf, kDecPrev, kDecPrev, kDecPrev, kDecPrev, kDecPrev, kDecPrev };
        for (int i = 1; i < 8; i++) stF(W.cost + 8 * (size_t)h0 + i, cl);
        stD(W.dec + 8 * (size_t)h0, d);
        stI(W.pcnt + h0, 1);
    }
    int cur = h0;
    while (cur != root) {
        const int par = W.parent[cur];
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
        __builtin_amdgcn_s_waitcnt(kWaitVmcnt0);   // this lane's tables performed before the arrival
        if (atomicAdd(&W.flag[par], 1) == 0) break;   // first to arrive: the sibling's lane computes the parent
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
        const int l = leftOf(hn, par), r = rightOf(hn, par);
        float cl[8], cr[8];
        for (int i = 1; i < 8; i++) { cl[i] = ldF(W.cost + 8 * (size_t)l + i); cr[i] = ldF(W.cost + 8 * (size_t)r + i); }
        const int p = ldI(W.pcnt + l) + ldI(W.pcnt + r);
        const float A = areaOf(boxLo(hn, par), boxHi(hn, par));
        float dist[9]; int dk[9];
        for (int j = 2; j <= 8; j++) {
            dist[j] = kInf; dk[j] = 1;
            for (int k = 1; k < j; k++) { const float v = cl[k] + cr[j - k]; if (v < dist[j]) { dist[j] = v; dk[j] = k; } }
        }
        const float cLeaf = p <= kPMax ? A * (float)p * kCPrim : kInf;
        const float cInt = dist[8] + A * kCNode;
        float c[8];
        unsigned char d[8];
        c[0] = 0.f;
        if (cLeaf <= cInt) { c[1] = cLeaf; d[1] = kDecLeaf; } else { c[1] = cInt; d[1] = kDecNode; }
        d[0] = (unsigned char)dk[8];
        for (int i = 2; i < 8; i++) {
            if (dist[i] < c[i - 1]) { c[i] = dist[i]; d[i] = (unsigned char)(kDecDist + dk[i]); }
            else { c[i] = c[i - 1]; d[i] = kDecPrev; }
        }
        for (int i = 1; i < 8; i++) stF(W.cost + 8 * (size_t)par + i, c[i]);
        stD(W.dec + 8 * (size_t)par, d);
        stI(W.pcnt + par, p);
        cur = par;
    }
}

// --------------------------------------------------------------------------------------------------- wide nodes
// the binary roots that node b as i wide roots stands for, in the host's recursion order (left subtree first)
__device__ int collectRoots(const Vec4* hn, int n, const unsigned char* dec, int b, int i, int* out, int m) {
    int stN[16], stI2[16], sp = 0;
    stN[sp] = b; stI2[sp] = i; sp++;
    while (sp > 0) {
        sp--;
        int nd = stN[sp], ii = stI2[sp];
        for (;;) {
            if (ii == 1 || nd < n) { out[m++] = nd; break; }   // a leaf's tables reduce every width to 1
            const unsigned char d = dec[8 * (size_t)nd + ii];
            if (d == kDecPrev) { ii--; continue; }
            const int k = d - kDecDist;
            stN[sp] = rightOf(hn, nd); stI2[sp] = ii - k; sp++;
            nd = leftOf(hn, nd); ii = k;
        }
    }
    return m;
}

// One wide node X of a level: children (decision roots), greedy octant slot assignment (repeatedly the (child, free
// slot) pair of lowest -dot(child centre - node centre, slot direction)), leaf flags, and the next level's records
// appended at ctl[1]. The root is a single leaf when its table chose a leaf (at most pMax primitives) or n == 1.
__device__ void cbLevelNode(const Vec4* __restrict__ hn, int n, int X, int cap, const Scratch& W) {
    const int b = W.wBin[X];
    const bool asLeaf = X == 0 && (n == 1 || W.dec[8 * (size_t)(2 * n - 2) + 1] == kDecLeaf);
    int ch[8], nc = 0;
    if (asLeaf) ch[nc++] = b;
    else {
        const int k8 = W.dec[8 * (size_t)b];
        nc = collectRoots(hn, n, W.dec, leftOf(hn, b), k8, ch, 0);
        nc = collectRoots(hn, n, W.dec, rightOf(hn, b), 8 - k8, ch, nc);
    }
    const Vec3 lo = boxLo(hn, b), hi = boxHi(hn, b);
    const Vec3 pc = (lo + hi) * 0.5f;
    float cost[8][8];
    bool used[8];
    int slotChild[8];
    for (int s = 0; s < 8; s++) slotChild[s] = -1;
    for (int c = 0; c < 8; c++) used[c] = false;
    for (int c = 0; c < nc; c++) {
        const Vec3 cc = (boxLo(hn, ch[c]) + boxHi(hn, ch[c])) * 0.5f - pc;
        for (int s = 0; s < 8; s++)
            cost[c][s] = -(cc.x * (s & 1 ? -1.f : 1.f) + cc.y * (s & 2 ? -1.f : 1.f) + cc.z * (s & 4 ? -1.f : 1.f));
    }
    for (int it = 0; it < nc; it++) {
        float bc = kInf; int bi = -1, bs = -1;
        for (int c = 0; c < nc; c++)
            if (!used[c])
                for (int s = 0; s < 8; s++)
                    if (slotChild[s] < 0 && cost[c][s] < bc) { bc = cost[c][s]; bi = c; bs = s; }
        if (bi < 0)
            for (int c = 0; c < nc && bi < 0; c++)
                if (!used[c])
                    for (int s = 0; s < 8; s++)
                        if (slotChild[s] < 0) { bi = c; bs = s; break; }
        used[bi] = true; slotChild[bs] = ch[bi];
    }
    unsigned leaf = 0;
    int nInt = 0, lp = 0, maxLeaf = 0;
    for (int s = 0; s < 8; s++) {
        const int cn = slotChild[s];
        W.wSlot[8 * (size_t)X + s] = cn;
        if (cn < 0) continue;
        const bool lf = asLeaf || cn < n || W.dec[8 * (size_t)cn + 1] == kDecLeaf;
        if (lf) { leaf |= 1u << s; lp += W.pcnt[cn]; maxLeaf = max(maxLeaf, W.pcnt[cn]); }
        else nInt++;
    }
    W.wLeaf[X] = (unsigned char)leaf; W.wNInt[X] = nInt; W.wLp[X] = lp;
    if (maxLeaf) atomicMax(&W.ctl[2], maxLeaf);
    if (nInt) {
        const int base = atomicAdd(&W.ctl[1], nInt);
        if (base + nInt > cap) { atomicOr(&W.ctl[0], 8); return; }
        int ord = 0;
        for (int s = 0; s < 8; s++) {
            const int cn = slotChild[s];
            if (cn < 0 || ((leaf >> s) & 1u)) continue;
            W.wBin[base + ord] = cn;
            W.wChild[8 * (size_t)X + ord] = base + ord;
            ord++;
        }
    }
}

// Level `l` of the wide tree, its range read from lv (the host does not wait for it): a grid-stride walk. Nothing runs
// after an error (ctl[0]), and the range is clamped to the wide-node capacity.
__global__ void __launch_bounds__(256) k_cb_level(const Vec4* __restrict__ hn, int n, int l, int cap, Scratch W) {
    if (W.ctl[0]) return;
    const int start = W.lv[l], end = min(W.lv[l + 1], cap);
    for (int X = start + blockIdx.x * blockDim.x + threadIdx.x; X < end; X += gridDim.x * blockDim.x)
        cbLevelNode(hn, n, X, cap, W);
}
// after level l: level l + 1 ends where the appends of level l ended
__global__ void k_cb_snap(int l, Scratch W) {
    if (threadIdx.x == 0 && blockIdx.x == 0) W.lv[l + 2] = W.ctl[1];
}

// subtree totals of internal children and leaf primitives, one level at a time from the deepest
__global__ void k_cb_totals(int start, int end, Scratch W) {
    const int X = start + blockIdx.x * blockDim.x + threadIdx.x;
    if (X >= end) return;
    int t = W.wNInt[X], tp = W.wLp[X];
    for (int j = 0; j < W.wNInt[X]; j++) { const int c = W.wChild[8 * (size_t)X + j]; t += W.wT[c]; tp += W.wTP[c]; }
    W.wT[X] = t; W.wTP[X] = tp;
}

// the host's numbering, one level at a time from the root: its stack pops a node, appends the node's internal
// children (child base) and leaf primitives (triangle base), pushes the internal children in slot order and so visits
// them in decreasing slot order
__global__ void k_cb_offsets(int start, int end, Scratch W) {
    const int X = start + blockIdx.x * blockDim.x + threadIdx.x;
    if (X >= end) return;
    const int m = W.wNInt[X];
    int s = W.wS[X] + m, p = W.wP[X] + W.wLp[X];
    for (int j = m - 1; j >= 0; j--) {   // later slots are visited first
        const int c = W.wChild[8 * (size_t)X + j];
        W.wS[c] = s; W.wP[c] = p;
        W.wIdx[c] = 1 + W.wS[X] + j;
        s += W.wT[c]; p += W.wTP[c];
    }
}

__device__ int quantExpD(float ext, float lo, float hi) {
    if (!(ext > 0.f)) return -100;
    int e = (int)ceilf((float)log2((double)(ext / 255.f)));   // the correctly rounded single-precision log2
    if (e < -100) e = -100;
    while (lo + 255.f * ldexpf(1.f, e) < hi) e++;
    return e;
}
__device__ int quantD(float v, float o, float sc, bool up) {
    const float f = (v - o) / sc;
    const int i = up ? (int)ceilf(f) : (int)floorf(f);
    return i < 0 ? 0 : (i > 255 ? 255 : i);
}
__device__ uint32_t pack4D(const int* q, int b) {
    return (uint32_t)q[b] | ((uint32_t)q[b + 1] << 8) | ((uint32_t)q[b + 2] << 16) | ((uint32_t)q[b + 3] << 24);
}

// the 80-byte node of wide node X and its leaves' primitives
__global__ void __launch_bounds__(256) k_cb_emit(const Vec4* __restrict__ hn, int n, int total, Scratch W, Vec4* out, int* prim) {
    const int X = blockIdx.x * blockDim.x + threadIdx.x;
    if (X >= total) return;
    const int b = W.wBin[X];
    const Vec3 lo = boxLo(hn, b), hi = boxHi(hn, b);
    const Vec3 ext = hi - lo;
    const int ex = quantExpD(ext.x, lo.x, hi.x), ey = quantExpD(ext.y, lo.y, hi.y), ez = quantExpD(ext.z, lo.z, hi.z);
    const float sx = ldexpf(1.f, ex), sy = ldexpf(1.f, ey), sz = ldexpf(1.f, ez);
    int qlox[8], qloy[8], qloz[8], qhix[8], qhiy[8], qhiz[8], meta[8];
    const int triBase = W.wP[X], childBase = 1 + W.wS[X];
    const unsigned leaf = W.wLeaf[X];
    unsigned imask = 0;
    int off = 0;
    for (int s = 0; s < 8; s++) {
        qlox[s] = qloy[s] = qloz[s] = qhix[s] = qhiy[s] = qhiz[s] = 0; meta[s] = 0;
        const int cn = W.wSlot[8 * (size_t)X + s];
        if (cn < 0) continue;
        const Vec3 cl = boxLo(hn, cn), chh = boxHi(hn, cn);
        qlox[s] = quantD(cl.x, lo.x, sx, false); qhix[s] = quantD(chh.x, lo.x, sx, true);
        qloy[s] = quantD(cl.y, lo.y, sy, false); qhiy[s] = quantD(chh.y, lo.y, sy, true);
        qloz[s] = quantD(cl.z, lo.z, sz, false); qhiz[s] = quantD(chh.z, lo.z, sz, true);
        if ((leaf >> s) & 1u) {   // the binary subtree's primitives, left to right
            int st[8], sp = 0, cnt = 0;
            st[sp++] = cn;
            while (sp > 0) {   // at most pMax leaves: the stack never holds more than pMax + 1 entries
                const int h = st[--sp];
                if (h < n) { if (triBase + off + cnt < n) prim[triBase + off + cnt] = leftOf(hn, h); cnt++; }
                else if (sp < 7) { st[sp++] = rightOf(hn, h); st[sp++] = leftOf(hn, h); }
            }
            meta[s] = 0x80 | ((cnt - 1) << 5) | off;
            off += cnt;
        } else { meta[s] = 0x40; imask |= 1u << s; }
    }
    const uint32_t ew = (uint32_t)(ex + 127) | ((uint32_t)(ey + 127) << 8) | ((uint32_t)(ez + 127) << 16) | (imask << 24);
    if (W.wIdx[X] >= total) return;   // the numbering covers [0, total) exactly
    Vec4* o = out + 5 * (size_t)W.wIdx[X];
    o[0] = v4(lo.x, lo.y, lo.z, asFloatU(ew));
    o[1] = v4(asFloat(childBase), asFloat(triBase), asFloatU(pack4D(meta, 0)), asFloatU(pack4D(meta, 4)));
    const uint32_t u[12] = { pack4D(qlox, 0), pack4D(qlox, 4), pack4D(qloy, 0), pack4D(qloy, 4), pack4D(qloz, 0), pack4D(qloz, 4),
                             pack4D(qhix, 0), pack4D(qhix, 4), pack4D(qhiy, 0), pack4D(qhiy, 4), pack4D(qhiz, 0), pack4D(qhiz, 4) };
    for (int j = 0; j < 3; j++) o[2 + j] = v4(asFloatU(u[4 * j]), asFloatU(u[4 * j + 1]), asFloatU(u[4 * j + 2]), asFloatU(u[4 * j + 3]));
}

__global__ void k_cb_root(int n, Scratch W) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    W.wBin[0] = 2 * n - 2; W.wS[0] = 0; W.wP[0] = 0; W.wIdx[0] = 0;
    W.ctl[1] = 1;
    W.lv[0] = 0; W.lv[1] = 1;
}

#define CCHK(x)                                                                                                        \
    do {                                                                                                               \
        hipError_t e_ = (x);                                                                                           \
        if (e_ != hipSuccess) { err = std::string(#x) + ": " + hipGetErrorString(e_); return false; }                  \
    } while (0)

float hostArea(Vec3 lo, Vec3 hi) {
    if (hi.x < lo.x) return 0.f;
    const Vec3 d = hi - lo;
    return 2.f * (d.x * d.y + d.y * d.z + d.z * d.x);
}

Scratch carve(void* mem, int n) {
    const size_t N = 2 * (size_t)n - 1, M = (size_t)n;
    char* p = (char*)mem;
    auto take = [&](size_t bytes) { char* q = p; p += (bytes + 255) & ~(size_t)255; return (void*)q; };
    Scratch W;
    W.ctl = (int*)take(64);
    W.parent = (int*)take(N * 4); W.pcount = (int*)take(N * 4); W.flag = (int*)take(N * 4); W.pcnt = (int*)take(N * 4);
    W.cost = (float*)take(N * 32); W.dec = (unsigned char*)take(N * 8);
    W.wBin = (int*)take(M * 4); W.wSlot = (int*)take(M * 32); W.wChild = (int*)take(M * 32); W.wNInt = (int*)take(M * 4);
    W.wLp = (int*)take(M * 4); W.wT = (int*)take(M * 4); W.wTP = (int*)take(M * 4); W.wS = (int*)take(M * 4);
    W.wP = (int*)take(M * 4); W.wIdx = (int*)take(M * 4); W.wLeaf = (unsigned char*)take(M);
    W.lv = (int*)take((M + 2 + kLevelBatch) * 4);   // snapshots run up to a batch past the last level
    return W;
}
size_t scratchBytes(int n) {
    const size_t N = 2 * (size_t)n - 1, M = (size_t)n;
    auto r = [](size_t b) { return (b + 255) & ~(size_t)255; };
    return r(64) + 4 * r(N * 4) + r(N * 32) + r(N * 8) + r(M * 4) + 2 * r(M * 32) + 7 * r(M * 4) + r(M) + r((M + 2 + kLevelBatch) * 4) + 256;
}

}  // namespace

CwbvhCollapser::~CwbvhCollapser() { release(); }

void CwbvhCollapser::release() {
    if (mem_) hipFree(mem_);
    mem_ = nullptr; cap_ = 0; bytes_ = 0;
}

bool CwbvhCollapser::reserve(int n, std::string& err) {
    if (n <= cap_) return true;
    release();
    bytes_ = scratchBytes(n);
    CCHK(hipMalloc(&mem_, bytes_));
    cap_ = n;
    return true;
}

bool CwbvhCollapser::run(const Vec4* hn, int n, Vec4* nodesOut, int* primOut, void* stream, int& numNodes, int& depth,
                         std::string& err) {
    if (n < 1 || !hn || !nodesOut || !primOut) { err = "CWBVH collapse: bad arguments"; return false; }
    if (n > (1 << 28)) { err = "CWBVH collapse: too many primitives"; return false; }
    if (!reserve(n, err)) return false;
    hipStream_t st = (hipStream_t)stream;
    const Scratch W = carve(mem_, n);
    const int N = 2 * n - 1;
    // zero the counters
    CCHK(hipMemsetAsync(W.ctl, 0, 64, st));
    CCHK(hipMemsetAsync(W.pcount, 0, (size_t)N * 4, st));
    CCHK(hipMemsetAsync(W.flag, 0, (size_t)N * 4, st));
    const int gN = std::min((N + 255) / 256, 4096);
    k_cb_check<<<gN, 256, 0, st>>>(hn, n, W);
    k_cb_check2<<<gN, 256, 0, st>>>(n, W);
    int ctl[3] = { 0, 0, 0 };
    CCHK(hipMemcpyAsync(ctl, W.ctl, 12, hipMemcpyDeviceToHost, st));
    CCHK(hipStreamSynchronize(st));
    if (ctl[0]) { err = "CWBVH collapse: the H-PLOC nodes are not a tree over [0, n) (code " + std::to_string(ctl[0]) + ")"; return false; }
    k_cb_dp<<<(n + 255) / 256, 256, 0, st>>>(hn, n, W);
    k_cb_root<<<1, 64, 0, st>>>(n, W);
    // The wide levels, kLevelBatch launches between host checks: each level reads its range from lv, and a snapshot
    // after it records where the next level ends. The loop stops once the next level is empty (at most n levels:
    // every non-empty level holds a wide node; levels after an empty one are empty and launch no work).
    const int gL = std::max(1, std::min((n + 255) / 256, 2048));
    int l = 0, lvPair[2] = { 0, 1 };
    while (lvPair[1] > lvPair[0]) {
        if (l >= n) { err = "CWBVH collapse: wide tree deeper than its primitive count"; return false; }
        for (int k = 0; k < kLevelBatch; k++, l++) {
            k_cb_level<<<gL, 256, 0, st>>>(hn, n, l, n, W);
            k_cb_snap<<<1, 64, 0, st>>>(l, W);
        }
        CCHK(hipMemcpyAsync(ctl, W.ctl, 12, hipMemcpyDeviceToHost, st));
        CCHK(hipStreamSynchronize(st));
        if (ctl[0]) { err = "CWBVH collapse: more wide nodes than primitives (code " + std::to_string(ctl[0]) + ")"; return false; }
        // levels 0 .. l - 1 are done; lv[l], lv[l + 1] bound level l (lv[l + 1]: the snapshot after level l - 1)
        CCHK(hipMemcpy(lvPair, W.lv + l, 8, hipMemcpyDeviceToHost));
    }
    // levels: the non-empty ones (a prefix of 0 .. l - 1); starts: their boundaries (the last entry is the total)
    std::vector<int> lvAll(l + 2);
    CCHK(hipMemcpy(lvAll.data(), W.lv, (size_t)(l + 2) * 4, hipMemcpyDeviceToHost));
    int levels = 0;
    while (levels <= l && lvAll[levels + 1] > lvAll[levels]) levels++;
    std::vector<int> starts(lvAll.begin(), lvAll.begin() + levels + 1);
    unsigned char rootDec[2] = { 0, 0 };
    CCHK(hipMemcpy(rootDec, W.dec + 8 * (size_t)(N - 1), 2, hipMemcpyDeviceToHost));
    CCHK(hipMemcpy(&rootCost_, W.cost + 8 * (size_t)(N - 1) + 1, 4, hipMemcpyDeviceToHost));
    rootLeaf_ = n == 1 || rootDec[1] == kDecLeaf;
    for (int l = levels - 1; l >= 0; l--)
        k_cb_totals<<<(starts[l + 1] - starts[l] + 255) / 256, 256, 0, st>>>(starts[l], starts[l + 1], W);
    for (int l = 0; l < levels; l++)
        k_cb_offsets<<<(starts[l + 1] - starts[l] + 255) / 256, 256, 0, st>>>(starts[l], starts[l + 1], W);
    const int total = starts.back();
    k_cb_emit<<<(total + 255) / 256, 256, 0, st>>>(hn, n, total, W, nodesOut, primOut);
    CCHK(hipGetLastError());
    numNodes = total; depth = levels;
    return true;
}

bool cwbvhFromHplocHost(const Vec4* hnodes, int n, std::vector<Vec4>& nodes, std::vector<int>& prim, int& depth,
                        double& gpuMs, float& sah, std::string& err) {
    if (n < 1) { err = "CWBVH collapse: no primitives"; return false; }
    const size_t N = 2 * (size_t)n - 1;
    Vec4 *dH = nullptr, *dN = nullptr;
    int* dP = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    CwbvhCollapser C;
    bool ok = false;
    do {
        if (hipMalloc(&dH, N * 2 * sizeof(Vec4)) != hipSuccess || hipMalloc(&dN, (size_t)n * 5 * sizeof(Vec4)) != hipSuccess ||
            hipMalloc(&dP, (size_t)n * 4) != hipSuccess) { err = "CWBVH collapse: out of device memory"; break; }
        if (hipMemcpy(dH, hnodes, N * 2 * sizeof(Vec4), hipMemcpyHostToDevice) != hipSuccess) { err = "CWBVH collapse: upload failed"; break; }
        if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) { err = "CWBVH collapse: events"; break; }
        int nn = 0;
        if (!C.run(dH, n, dN, dP, nullptr, nn, depth, err)) break;   // first run: loads the kernels (not timed)
        hipEventRecord(e0, nullptr);
        if (!C.run(dH, n, dN, dP, nullptr, nn, depth, err)) break;
        hipEventRecord(e1, nullptr);
        if (hipEventSynchronize(e1) != hipSuccess) { err = "CWBVH collapse failed"; break; }
        float ms = 0;
        hipEventElapsedTime(&ms, e0, e1);
        gpuMs = ms;
        {   // the host collapse's SAH: the root's cost (an internal wide root; or a leaf root plus one node) over its area
            const Vec4 rl = hnodes[2 * (2 * (size_t)n - 2)], rh = hnodes[2 * (2 * (size_t)n - 2) + 1];
            const float Ar = hostArea(xyz(rl), xyz(rh));
            const bool leafRoot = n == 1 || C.rootIsLeaf();
            sah = Ar > 0 ? (leafRoot ? (Ar * (float)n * kCPrim + Ar * kCNode) / Ar : C.rootCost() / Ar) : 0.f;
        }
        nodes.resize((size_t)nn * 5); prim.resize((size_t)n);
        if (hipMemcpy(nodes.data(), dN, nodes.size() * sizeof(Vec4), hipMemcpyDeviceToHost) != hipSuccess ||
            hipMemcpy(prim.data(), dP, prim.size() * 4, hipMemcpyDeviceToHost) != hipSuccess) { err = "CWBVH collapse: download failed"; break; }
        ok = true;
    } while (false);
    if (e0) hipEventDestroy(e0);
    if (e1) hipEventDestroy(e1);
    if (dH) hipFree(dH);
    if (dN) hipFree(dN);
    if (dP) hipFree(dP);
    return ok;
}

}  // namespace cwp
