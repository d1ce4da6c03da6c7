// device/hploc.hip - H-PLOC BVH2 builder for gfx950 (see hploc.h). Written from the H-PLOC paper (Benthin et al.
// 2024) and the reference's description of its variant (src/accel/hploc.h:1-11): Morton-keyed LBVH climbed bottom-up
// (Apetrei 2014), PLOC merges of 2 x 16 clusters within a radius of 8 positions.
#include "device/hploc.h"
#include <mutex>
#include <hip/hip_runtime.h>
#include <algorithm>

namespace cwp {

namespace {

constexpr int kT = 16;          // clusters kept per LBVH subtree
constexpr int kR = 8;           // PLOC nearest-neighbour radius (positions in Morton order)
constexpr int kSortBlock = 256; // radix sort workgroup: 4 waves, one thread per digit value
constexpr int kSortItems = 16;  // keys per thread per pass
constexpr int kTile = kSortBlock * kSortItems;
constexpr int kSmallSort = 16384;   // up to this many keys are sorted by one workgroup (no inter-workgroup passes)
constexpr int kRankSort = 4096;     // up to this many (unique) keys are placed by their rank in one launch

// order-preserving float <-> uint mapping, for min / max through integer atomics
__device__ __forceinline__ uint32_t f2o(float f) {
    const uint32_t u = __float_as_uint(f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float o2f(uint32_t u) { return __uint_as_float((u & 0x80000000u) ? (u & 0x7fffffffu) : ~u); }

// spreads the low 21 bits of v to every third bit of a 63-bit word (bit b -> bit 3 b)
__device__ __forceinline__ uint64_t spread3(uint32_t v) {
    uint64_t x = v & 0x1fffffu;
    x = (x | (x << 32)) & 0x001f00000000ffffull;
    x = (x | (x << 16)) & 0x001f0000ff0000ffull;
    x = (x | (x << 8)) & 0x100f00f00f00f00full;
    x = (x | (x << 4)) & 0x10c30c30c30c30c3ull;
    x = (x | (x << 2)) & 0x1249249249249249ull;
    return x;
}

__device__ __forceinline__ int lanePrefix(uint64_t mask) {   // set bits of mask below this lane
    return (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
}

// min / max as comparisons (not fminf / fmaxf), so signed zeros resolve the same way on every machine
__device__ __forceinline__ float mn(float a, float b) { return a < b ? a : b; }
__device__ __forceinline__ float mx(float a, float b) { return a > b ? a : b; }
__device__ __forceinline__ Vec3 vmn(Vec3 a, Vec3 b) { return v3(mn(a.x, b.x), mn(a.y, b.y), mn(a.z, b.z)); }
__device__ __forceinline__ Vec3 vmx(Vec3 a, Vec3 b) { return v3(mx(a.x, b.x), mx(a.y, b.y), mx(a.z, b.z)); }
__device__ __forceinline__ float halfArea(Vec3 lo, Vec3 hi) {
    const float dx = hi.x - lo.x, dy = hi.y - lo.y, dz = hi.z - lo.z;
    return dx * dy + dy * dz + dz * dx;
}
__device__ __forceinline__ Vec3 shfl3(Vec3 v, int src) { return v3(__shfl(v.x, src), __shfl(v.y, src), __shfl(v.z, src)); }

// Data handed between waves of the climb (cluster ids, node boxes) is read and written with relaxed device-scope
// atomic accesses: on gfx950 these carry the device-coherence bit, so a wave on one XCD sees what a wave on another
// XCD wrote once the hand-over atomic orders them, without the whole-L2 write-back and invalidate of an agent-scope
// fence (the L2s of the 8 XCDs are not coherent with each other). Workgroup-scope fences keep the compiler and the
// memory counters in order around the hand-over.
constexpr int kWaitVmcnt0 = 0x0F70;   // s_waitcnt vmcnt(0) (expcnt and lgkmcnt left at their maxima)
__device__ __forceinline__ int ldC(const int* p) { return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); }
__device__ __forceinline__ void stC(int* p, int v) { __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); }
__device__ __forceinline__ Vec3 ldBox3(const Vec4* p) {   // xyz of a node record
    const uint64_t a = __hip_atomic_load((const uint64_t*)p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const uint32_t z = __hip_atomic_load((const uint32_t*)p + 2, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    return v3(__uint_as_float((uint32_t)a), __uint_as_float((uint32_t)(a >> 32)), __uint_as_float(z));
}
__device__ __forceinline__ Vec4 ldBox4(const Vec4* p) {   // a cluster box record with its w component
    const uint64_t a = __hip_atomic_load((const uint64_t*)p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const uint64_t b = __hip_atomic_load((const uint64_t*)p + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    return v4(__uint_as_float((uint32_t)a), __uint_as_float((uint32_t)(a >> 32)), __uint_as_float((uint32_t)b),
              __uint_as_float((uint32_t)(b >> 32)));
}
__device__ __forceinline__ void stNode(Vec4* p, Vec4 v) {
    const uint64_t a = (uint64_t)__float_as_uint(v.x) | ((uint64_t)__float_as_uint(v.y) << 32);
    const uint64_t b = (uint64_t)__float_as_uint(v.z) | ((uint64_t)__float_as_uint(v.w) << 32);
    __hip_atomic_store((uint64_t*)p, a, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_store((uint64_t*)p + 1, b, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// ------------------------------------------------------------------------------------------------ keys
// build state reset; the grid also marks every temporary slot empty (height -1)
__global__ void k_hp_init(uint32_t* cbox, int* counter, int* status, int* theight, int tcap) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < 3) { cbox[t] = 0xffffffffu; cbox[3 + t] = 0u; }
    if (t == 0) { *counter = 0; status[0] = 0; status[1] = 0; }
    if (t < tcap) theight[t] = -1;
}

__global__ void __launch_bounds__(256) k_hp_bounds(const Vec4* __restrict__ lo, const Vec4* __restrict__ hi, int n, uint32_t* cbox) {
    __shared__ uint32_t red[4][6];
    uint32_t a[3] = { 0xffffffffu, 0xffffffffu, 0xffffffffu }, b[3] = { 0u, 0u, 0u };
    // four boxes per thread per round: their loads are issued together (the rounds are latency-bound; more
    // workgroups instead would put more atomics on the six bounds)
    const int S = gridDim.x * blockDim.x;
    for (int i0 = blockIdx.x * blockDim.x + threadIdx.x; i0 < n; i0 += 4 * S) {
        Vec4 l[4], h[4];
#pragma unroll
        for (int u = 0; u < 4; u++) {
            const int i = i0 + u * S;
            if (i < n) { l[u] = lo[i]; h[u] = hi[i]; }
        }
#pragma unroll
        for (int u = 0; u < 4; u++) {
            if (i0 + u * S >= n) break;
            const uint32_t c[3] = { f2o((l[u].x + h[u].x) * 0.5f), f2o((l[u].y + h[u].y) * 0.5f), f2o((l[u].z + h[u].z) * 0.5f) };
            for (int k = 0; k < 3; k++) { a[k] = min(a[k], c[k]); b[k] = max(b[k], c[k]); }
        }
    }
    for (int k = 0; k < 3; k++)
        for (int off = 32; off > 0; off >>= 1) {
            a[k] = min(a[k], (uint32_t)__shfl_xor((int)a[k], off));
            b[k] = max(b[k], (uint32_t)__shfl_xor((int)b[k], off));
        }
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0)
        for (int k = 0; k < 3; k++) { red[w][k] = a[k]; red[w][3 + k] = b[k]; }
    __syncthreads();
    if (threadIdx.x < 6) {   // one atomic per workgroup and bound
        const int k = threadIdx.x;
        uint32_t v = red[0][k];
        for (int q = 1; q < (int)(blockDim.x >> 6); q++) v = k < 3 ? min(v, red[q][k]) : max(v, red[q][k]);
        if (k < 3) atomicMin(&cbox[k], v); else atomicMax(&cbox[k], v);
    }
}

// key = Morton code of the box centroid (mb bits per axis, x << 2 | y << 1 | z) << ib | primitive index
__global__ void k_hp_keys(const Vec4* __restrict__ lo, const Vec4* __restrict__ hi, int n, const uint32_t* cbox,
                          uint64_t* keys, int mb, int ib) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float bl[3] = { o2f(cbox[0]), o2f(cbox[1]), o2f(cbox[2]) }, bh[3] = { o2f(cbox[3]), o2f(cbox[4]), o2f(cbox[5]) };
    const Vec4 l = lo[i], h = hi[i];
    const float c[3] = { (l.x + h.x) * 0.5f, (l.y + h.y) * 0.5f, (l.z + h.z) * 0.5f };
    const float s = (float)(1u << mb), smax = s - 1.f;
    uint32_t q[3];
    for (int k = 0; k < 3; k++) {
        const float e = bh[k] - bl[k];
        const float t = e > 0.f ? (c[k] - bl[k]) / e : 0.f;
        float v = t * s;
        v = v > 0.f ? v : 0.f;
        v = v < smax ? v : smax;
        q[k] = (uint32_t)v;
    }
    const uint64_t m = (spread3(q[0]) << 2) | (spread3(q[1]) << 1) | spread3(q[2]);
    keys[i] = (m << ib) | (uint64_t)i;
}

// ------------------------------------------------------------------------------------------------ radix sort
// Stable ranking of one 256-key round inside a 4-wave workgroup: each lane gets the number of keys with the same
// digit before it in its wave (8 ballots over the digit bits) and the wave's count per digit goes to wh[wave][digit].
__device__ __forceinline__ int rankInWave(bool valid, uint32_t d, int* whRow) {
    uint64_t m = __ballot(valid);
    for (int b = 0; b < 8; b++) {
        const uint64_t bb = __ballot((d >> b) & 1u);
        m &= ((d >> b) & 1u) ? bb : ~bb;
    }
    const int rank = lanePrefix(m);
    if (valid && rank == 0) whRow[d] = (int)__builtin_popcountll(m);
    return rank;
}

// exclusive prefix sum over the 256 threads of a workgroup (tmp: 4 ints of LDS); returns this thread's prefix
__device__ __forceinline__ int blockScan256(int v, int* tmp, int& total) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    int x = v;
    for (int off = 1; off < 64; off <<= 1) {
        const int y = __shfl_up(x, off);
        if (lane >= off) x += y;
    }
    if (lane == 63) tmp[w] = x;
    __syncthreads();
    int before = 0;
    for (int k = 0; k < w; k++) before += tmp[k];
    total = tmp[0] + tmp[1] + tmp[2] + tmp[3];
    __syncthreads();
    return before + x - v;
}

// digit histogram of one tile: counts[digit * tiles + tile]
// gate (optional): the pass runs only if *gate is non-zero (a pass whose need is known only on the device)
__global__ void __launch_bounds__(kSortBlock) k_rs_count(const uint64_t* __restrict__ keys, int n, int shift, int* counts, int tiles,
                                                         const int* gate = nullptr) {
    if (gate && *gate == 0) return;
    __shared__ int hist[256];
    hist[threadIdx.x] = 0;
    __syncthreads();
    const int base = blockIdx.x * kTile;
    for (int j = 0; j < kSortItems; j++) {
        const int idx = base + j * kSortBlock + (int)threadIdx.x;
        if (idx < n) atomicAdd(&hist[(uint32_t)(keys[idx] >> shift) & 255u], 1);
    }
    __syncthreads();
    counts[threadIdx.x * tiles + blockIdx.x] = hist[threadIdx.x];
}

// per digit (one workgroup per digit row of counts): exclusive scan of the row's tile counts in place, row total to
// totals[digit]
__global__ void __launch_bounds__(kSortBlock) k_rs_rowscan(int* counts, int tiles, int* totals, const int* gate = nullptr) {
    if (gate && *gate == 0) return;
    __shared__ int tmp[4];
    int* row = counts + (size_t)blockIdx.x * tiles;
    int carry = 0;
    for (int b = 0; b < tiles; b += kSortBlock) {
        const int j = b + (int)threadIdx.x;
        const int v = j < tiles ? row[j] : 0;
        int total;
        const int pre = blockScan256(v, tmp, total);
        if (j < tiles) row[j] = carry + pre;
        carry += total;
    }
    if (threadIdx.x == 0) totals[blockIdx.x] = carry;
}

// stable scatter of one tile by the digit at `shift`: the tile's start for digit d is the exclusive scan of the digit
// totals (computed here by every workgroup) plus the row prefix of k_rs_rowscan
__global__ void __launch_bounds__(kSortBlock) k_rs_scatter(const uint64_t* __restrict__ in, uint64_t* __restrict__ out, int n,
                                                           int shift, const int* __restrict__ offsets, int tiles,
                                                           const int* __restrict__ totals, const int* gate = nullptr) {
    if (gate && *gate == 0) return;
    __shared__ int base[256], running[256], wh[4][256], tmp[4];
    const int t = threadIdx.x, w = t >> 6;
    int all;
    const int digitStart = blockScan256(totals[t], tmp, all);
    base[t] = digitStart + offsets[t * tiles + blockIdx.x];
    running[t] = 0;
    const int tile = blockIdx.x * kTile;
    for (int j = 0; j < kSortItems; j++) {
        const int idx = tile + j * kSortBlock + t;
        const bool valid = idx < n;
        const uint64_t key = valid ? in[idx] : 0ull;
        const uint32_t d = (uint32_t)(key >> shift) & 255u;
        wh[0][t] = 0; wh[1][t] = 0; wh[2][t] = 0; wh[3][t] = 0;
        __syncthreads();
        const int rank = rankInWave(valid, d, wh[w]);
        __syncthreads();
        if (valid) {
            int pos = base[d] + running[d] + rank;
            for (int k = 0; k < w; k++) pos += wh[k][d];
            out[pos] = key;
        }
        __syncthreads();
        running[t] += wh[0][t] + wh[1][t] + wh[2][t] + wh[3][t];
        __syncthreads();
    }
}

// Sort of at most kRankSort unique keys in one launch: every key's position is the number of keys below it. Each
// workgroup ranks its 256 keys against all n, staged through LDS 256 at a time (every lane reads the same entry:
// broadcasts). Padding entries are ~0, above every key (keys use at most 63 bits).
__global__ void __launch_bounds__(kSortBlock) k_rank_sort(const uint64_t* __restrict__ in, uint64_t* __restrict__ out, int n) {
    __shared__ uint64_t tile[kSortBlock];
    const int i = blockIdx.x * kSortBlock + (int)threadIdx.x;
    const uint64_t key = i < n ? in[i] : ~0ull;
    int rank = 0;
    for (int b = 0; b < n; b += kSortBlock) {
        const int j = b + (int)threadIdx.x;
        tile[threadIdx.x] = j < n ? in[j] : ~0ull;
        __syncthreads();
#pragma unroll 8
        for (int k = 0; k < kSortBlock; k++) rank += tile[k] < key ? 1 : 0;
        __syncthreads();
    }
    if (i < n) out[rank] = key;
}

// whole sort in one workgroup (small n): every pass (digits at shift0, shift0 + 8, ...) counts, scans and scatters
// a <-> b; the result ends in a if `passes` is even, else in b
__global__ void __launch_bounds__(kSortBlock) k_rs_single(uint64_t* a, uint64_t* b, int n, int shift0, int passes) {
    __shared__ int hist[256], running[256], wh[4][256], tmp[4];
    const int t = threadIdx.x, w = t >> 6;
    uint64_t* src = a;
    uint64_t* dst = b;
    for (int p = 0; p < passes; p++) {
        const int shift = shift0 + 8 * p;
        hist[t] = 0;
        __syncthreads();
        for (int idx = t; idx < n; idx += kSortBlock) atomicAdd(&hist[(uint32_t)(src[idx] >> shift) & 255u], 1);
        __syncthreads();
        int total;
        const int start = blockScan256(hist[t], tmp, total);
        __syncthreads();
        hist[t] = start;
        running[t] = 0;
        __syncthreads();
        for (int r = 0; r < n; r += kSortBlock) {
            const int idx = r + t;
            const bool valid = idx < n;
            const uint64_t key = valid ? src[idx] : 0ull;
            const uint32_t d = (uint32_t)(key >> shift) & 255u;
            wh[0][t] = 0; wh[1][t] = 0; wh[2][t] = 0; wh[3][t] = 0;
            __syncthreads();
            const int rank = rankInWave(valid, d, wh[w]);
            __syncthreads();
            if (valid) {
                int pos = hist[d] + running[d] + rank;
                for (int k = 0; k < w; k++) pos += wh[k][d];
                dst[pos] = key;
            }
            __syncthreads();
            running[t] += wh[0][t] + wh[1][t] + wh[2][t] + wh[3][t];
            __syncthreads();
        }
        __threadfence_block();
        uint64_t* x = src; src = dst; dst = x;
        __syncthreads();
    }
}

// ------------------------------------------------------------------------------------------------ H-PLOC
__device__ __forceinline__ void hpLeaf(const Vec4* __restrict__ lo, const Vec4* __restrict__ hi, const uint64_t* __restrict__ keys,
                                       int i, int ib, Vec4* nodes, int* C, Vec4* CB, int* bounds) {
    const int p = (int)(keys[i] & ((1ull << ib) - 1ull));
    const Vec4 l = lo[p], h = hi[p];
    nodes[2 * i] = v4(xyz(l), asFloat(p));
    nodes[2 * i + 1] = v4(xyz(h), asFloat(-1));
    C[i] = i;
    CB[2 * i] = v4(xyz(l), 0.f);
    CB[2 * i + 1] = v4(xyz(h), 0.f);
    bounds[i] = 0;   // arrival counter of the LBVH node split at i
}

// Per half-wave LDS of a merge: compaction buffers (ids, boxes) and the nodes created by the merge, which get their
// global ids only when the merge is done (one counter reservation per merge instead of one per iteration).
struct MergeLds {
    int id[32], ht[32];
    float box[32 * 6];
    float node[32 * 8];   // lo xyz, hi xyz, left id bits, right id bits
    int nodeHt[32];
};

// Build-time node ids: leaves keep [0, n); a merge's nodes get temporary slots n + t from its wave's chunk of the
// temporary array (chunks of kChunk slots from one counter, so the counter sees one atomic per chunk instead of one
// per merge); every node carries its height (leaves 0), and a renumbering pass after the climb orders the internal
// nodes by height, which keeps every child below its parent and puts the root, the one node of largest height, last.
constexpr int kChunk = 64;
struct Chunk { int next, end; };   // wave-uniform

// PLOC on the clusters of LBVH node [L, R] split at s, run by one half-wave (32 lanes; both halves call together, a
// half with on = false idles). Its clusters sit at the start of each child's range in C (ids) and CB (boxes, height in
// the first record's w) (-1 = empty slot): up to 16 from each side, compacted into lanes in order. Each iteration:
// every lane finds the pair of smallest merged half-area among the lanes within kR of it (pairs ordered by (area bits,
// lower lane, upper lane), so the global best pair is mutual and every iteration merges); mutual pairs merge into a
// new node (lower lane = left child), numbered locally (-2 - k) in lane order; survivors are compacted in lane order.
// Stops at kT clusters (1 at the root). Then the merge's nodes take consecutive temporary slots of the wave's chunk
// and the survivors go back to the start of [L, R].
__device__ void plocMerge(bool on, int L, int R, int s, bool root, int n, Vec4* tnodes, int* theight, int tcap, int* C,
                          Vec4* CB, int* counter, Chunk& ck, MergeLds& sh) {
    const int lane = (int)threadIdx.x & 63, hl = lane & 31, hb = lane & 32;
    const uint64_t ltMask = (1ull << lane) - 1ull, halfMask = 0xffffffffull << hb;
    int nL = 0, nR = 0;
    if (on) { nL = min(s - L + 1, kT); nR = min(R - s, kT); }
    const int slot = hl < nL ? L + hl : (hl < nL + nR ? s + 1 + (hl - nL) : -1);
    int id = -1, ht = 0;
    Vec3 lo = v3(0.f, 0.f, 0.f), hi = lo;
    if (slot >= 0) {   // id, box and height of the slot, read together
        id = ldC(C + slot);
        const Vec4 b0 = ldBox4(CB + 2 * slot);
        lo = xyz(b0); ht = __float_as_int(b0.w);
        hi = ldBox3(CB + 2 * slot + 1);
    }
    const uint64_t vm = __ballot(id >= 0) & halfMask;
    int num = (int)__builtin_popcountll(vm);
    if (id >= 0) {
        const int dd = (int)__builtin_popcountll(vm & ltMask);
        sh.id[dd] = id; sh.ht[dd] = ht;
        float* b = sh.box + 6 * dd;
        b[0] = lo.x; b[1] = lo.y; b[2] = lo.z; b[3] = hi.x; b[4] = hi.y; b[5] = hi.z;
    }
    __syncthreads();
    id = -1;
    if (hl < num) {
        id = sh.id[hl]; ht = sh.ht[hl];
        const float* b = sh.box + 6 * hl;
        lo = v3(b[0], b[1], b[2]); hi = v3(b[3], b[4], b[5]);
    }
    __syncthreads();
    const int target = root ? 1 : kT;
    int made = 0;   // nodes this half created so far (uniform per half)
    while (__ballot(num > target)) {
        const bool act = num > target;
        uint64_t best = ~0ull;
        // not unrolled further: the fully unrolled search holds every neighbour's box at once (102 registers, 4 waves
        // per SIMD); two radii per trip keep the kernel at 8 waves
#pragma unroll 2
        for (int r = 1; r <= kR; r++) {
            const Vec3 olo = shfl3(lo, hb + min(hl + r, 31)), ohi = shfl3(hi, hb + min(hl + r, 31));
            uint64_t key = ~0ull;
            if (hl + r < num)
                key = ((uint64_t)__float_as_uint(halfArea(vmn(lo, olo), vmx(hi, ohi))) << 32) | ((uint64_t)hl << 5) | (uint64_t)(hl + r);
            if (key < best) best = key;
            const uint64_t k2 = __shfl(key, hb + max(hl - r, 0));
            if (hl >= r && k2 < best) best = k2;
        }
        const int pa = (int)((best >> 5) & 31u), pb = (int)(best & 31u);
        const int nn = hl < num ? (pa == hl ? pb : pa) : hl;
        const int nnn = __shfl(nn, hb + nn);
        const bool mutual = act && hl < num && nn != hl && nnn == hl;
        const bool lead = mutual && hl < nn;
        const Vec3 plo = shfl3(lo, hb + nn), phi = shfl3(hi, hb + nn);
        const int pid = __shfl(id, hb + nn), pht = __shfl(ht, hb + nn);
        const uint64_t lm = __ballot(lead) & halfMask;
        if (lead) {
            const int k = made + (int)__builtin_popcountll(lm & ltMask);
            lo = vmn(lo, plo); hi = vmx(hi, phi);
            ht = 1 + max(ht, pht);
            float* q = sh.node + 8 * k;
            q[0] = lo.x; q[1] = lo.y; q[2] = lo.z; q[3] = hi.x; q[4] = hi.y; q[5] = hi.z;
            q[6] = __int_as_float(id); q[7] = __int_as_float(pid);
            sh.nodeHt[k] = ht;
            id = -2 - k;
        }
        made += (int)__builtin_popcountll(lm);
        const bool keep = act && hl < num && !(mutual && hl > nn);
        const uint64_t km = __ballot(keep) & halfMask;
        if (keep) {
            const int dd = (int)__builtin_popcountll(km & ltMask);
            sh.id[dd] = id; sh.ht[dd] = ht;
            float* b = sh.box + 6 * dd;
            b[0] = lo.x; b[1] = lo.y; b[2] = lo.z; b[3] = hi.x; b[4] = hi.y; b[5] = hi.z;
        }
        __syncthreads();
        if (act) {
            num = (int)__builtin_popcountll(km);
            if (hl < num) {
                id = sh.id[hl]; ht = sh.ht[hl];
                const float* b = sh.box + 6 * hl;
                lo = v3(b[0], b[1], b[2]); hi = v3(b[3], b[4], b[5]);
            } else id = -1;
        }
        __syncthreads();
    }
    // temporary slots for both halves' nodes: the rest of the wave's chunk, then a new chunk (kChunk >= 2 x 31 covers
    // the remainder of any merge); slots follow the halves' local order
    const int made0 = __shfl(made, 0), made1 = __shfl(made, 32), total = made0 + made1;
    const int a0 = ck.next, r0 = min(total, ck.end - ck.next);
    int a1 = 0;
    if (total > r0) {
        int b = 0;
        if (lane == 0) b = atomicAdd(counter, kChunk);
        b = __shfl(b, 0);
        a1 = b; ck.end = b + kChunk; ck.next = b + (total - r0);
    } else ck.next += total;
    const int first = hb ? made0 : 0;
    auto slotOf = [&](int k) { const int g = first + k; return g < r0 ? a0 + g : a1 + (g - r0); };
    auto ref = [&](int x) { return x >= -1 ? x : n + slotOf(-2 - x); };   // leaf id, n + slot, or -1
    if (hl < made && slotOf(hl) < tcap) {   // node records are read only after the climb: plain stores
        const int t = slotOf(hl);
        const float* q = sh.node + 8 * hl;
        tnodes[2 * (size_t)t] = v4(q[0], q[1], q[2], asFloat(ref(__float_as_int(q[6]))));
        tnodes[2 * (size_t)t + 1] = v4(q[3], q[4], q[5], asFloat(ref(__float_as_int(q[7]))));
        theight[t] = sh.nodeHt[hl];
    }
    if (slot >= 0) stC(C + slot, -1);
    __syncthreads();
    if (on && hl < num) {
        const int g = ref(id);
        stC(C + L + hl, g - n < tcap ? g : -1);
        stNode(CB + 2 * (L + hl), v4(lo, __int_as_float(ht)));
        stNode(CB + 2 * (L + hl) + 1, v4(hi, 0.f));
    }
    __syncthreads();
}

// The LBVH over the sorted keys as a binary radix tree (Karras 2012): internal node i finds its range [min(i, j),
// max(i, j)] and its split g by binary searches on the common-prefix length of its keys (keys are unique). The
// node is stored under its split position g (the bottom-up climb's hand-over position: the range is the union of
// [L, g] and [g + 1, R]); the tree is the one the bottom-up rule builds (extend towards the neighbour with the
// longer common prefix).
__device__ __forceinline__ int prefixLen(const uint64_t* __restrict__ keys, int n, int i, int j) {
    return (j < 0 || j > n - 1) ? -1 : (int)__builtin_clzll(keys[i] ^ keys[j]);
}
__device__ __forceinline__ void lbvhRange(const uint64_t* __restrict__ keys, int n, int i, int* rangeL, int* rangeR) {
    const int d = prefixLen(keys, n, i, i + 1) - prefixLen(keys, n, i, i - 1) > 0 ? 1 : -1;
    const int dmin = prefixLen(keys, n, i, i - d);
    int lmax = 2;
    while (prefixLen(keys, n, i, i + lmax * d) > dmin) lmax *= 2;
    int l = 0;
    for (int t = lmax / 2; t >= 1; t /= 2)
        if (prefixLen(keys, n, i, i + (l + t) * d) > dmin) l += t;
    const int j = i + l * d;
    const int dnode = prefixLen(keys, n, i, j);
    int s = 0, t = l;
    do {
        t = (t + 1) >> 1;
        if (prefixLen(keys, n, i, i + (s + t) * d) > dnode) s += t;
    } while (t > 1);
    const int g = i + s * d + min(d, 0);
    rangeL[g] = min(i, j); rangeR[g] = max(i, j);
}

// Leaf records (nodes, cluster slots, arrival counters) and the LBVH ranges in one launch: the leaves' gathers and
// stores and the ranges' chains of dependent key loads overlap.
__global__ void k_hp_leaves_ranges(const Vec4* __restrict__ lo, const Vec4* __restrict__ hi, const uint64_t* __restrict__ keys,
                                   int n, int ib, Vec4* nodes, int* C, Vec4* CB, int* bounds, int* rangeL, int* rangeR) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) hpLeaf(lo, hi, keys, i, ib, nodes, C, CB, bounds);
    if (i < n - 1) lbvhRange(keys, n, i, rangeL, rangeR);
}

// H-PLOC over the LBVH nodes that merge: those whose range holds more than kT clusters, and the root. One lane per
// split position; a lane starts at a merge node none of whose children merge (its clusters are the leaves of its
// range), merges it, then climbs to the parent (extend towards the neighbour with the longer common prefix). A parent
// with two merging children is handed over through an arrival counter: the first lane stops, the second merges it.
// Nodes whose ranges hold at most kT clusters need no work and no hand-over. The wave runs the merges of its lanes two
// at a time (one per half-wave). Cluster and node records cross waves (and XCDs) through device-coherent atomic
// accesses; a vmcnt(0) wait orders them before the arrival.
__global__ void __launch_bounds__(64) k_hploc(Vec4* tnodes, int* theight, int tcap, const uint64_t* __restrict__ keys,
                                              const int* __restrict__ rangeL, const int* __restrict__ rangeR, int* C,
                                              Vec4* CB, int* arrive, int* counter, int n) {
    __shared__ MergeLds sh[2];
    Chunk ck{ 0, 0 };
    const int lane = (int)threadIdx.x, half = lane >> 5;
    const int g = blockIdx.x * 64 + lane;
    int s = g, L = 0, R = 0;
    bool active = false;
    if (g < n - 1) {
        L = rangeL[g]; R = rangeR[g];
        const bool root = L == 0 && R == n - 1;
        active = (R - L + 1 > kT || root) && g - L + 1 <= kT && R - g <= kT;   // merges, no merging child
    }
    while (__ballot(active)) {
        uint64_t mm = __ballot(active);
        while (mm) {
            const int a = __builtin_ctzll(mm);
            mm &= mm - 1;
            const int b = mm ? __builtin_ctzll(mm) : -1;
            if (mm) mm &= mm - 1;
            const int src = half == 0 ? a : b;
            const int sl = src < 0 ? 0 : src;
            const int l = __shfl(L, sl), r = __shfl(R, sl), sp = __shfl(s, sl);
            plocMerge(src >= 0, l, r, sp, l == 0 && r == n - 1, n, tnodes, theight, tcap, C, CB, counter, ck, sh[half]);
        }
        // this wave's merges performed (vmcnt 0) at the coherence point before the arrivals below
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
        __builtin_amdgcn_s_waitcnt(kWaitVmcnt0);
        if (active) {
            if (L == 0 && R == n - 1) active = false;   // the root is done
            else {
                const int p = (L == 0 || (R != n - 1 && (keys[R] ^ keys[R + 1]) < (keys[L - 1] ^ keys[L]))) ? R : L - 1;
                const int pl = rangeL[p], pr = rangeR[p];
                const int need = (p - pl + 1 > kT ? 1 : 0) + (pr - p > kT ? 1 : 0);   // merging children of p
                if (need == 2 && atomicAdd(&arrive[p], 1) == 0) active = false;   // the sibling's lane merges p
                s = p; L = pl; R = pr;
            }
        }
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");   // the sibling's clusters after its arrival
    }
}

// ------------------------------------------------------------------------------------------------ renumbering
// sort keys of the temporary slots: height (holes: 0xffff) above the slot index
__global__ void k_rn_keys(const int* __restrict__ theight, int used, uint64_t* keys, int* tall) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= used) return;
    const int h = theight[t];
    if (h >= 255) atomicOr(tall, 1);   // the low height byte alone no longer orders the nodes (holes are 0xffff)
    keys[t] = ((uint64_t)(uint32_t)(h < 0 ? 0xffff : min(h, 0xfffe)) << 32) | (uint64_t)(uint32_t)t;
}
// final ids: the p-th slot in height order becomes node n + p; the n - 1 nodes must be exactly the non-hole slots.
// The order is in `one` after the pass over the low height byte, in `two` when the high byte needed a pass too.
__global__ void k_rn_map(const uint64_t* __restrict__ one, const uint64_t* __restrict__ two, const int* tall, int used,
                         int n, int* newId, int* status) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= used) return;
    const uint64_t* sorted = *tall ? two : one;
    const uint32_t h = (uint32_t)(sorted[p] >> 32), t = (uint32_t)sorted[p];
    const bool hole = h == 0xffff;
    if (hole != (p >= n - 1)) atomicOr(status, 1);
    if (!hole && p < n - 1) newId[t] = n + p;
}
__global__ void k_rn_write(const Vec4* __restrict__ tnodes, const int* __restrict__ theight, const int* __restrict__ newId,
                           int used, int n, Vec4* nodes, int* status) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= used || theight[t] < 0) return;
    const int f = newId[t];
    if (f < n || f > 2 * n - 2) { atomicOr(status, 2); return; }
    Vec4 a = tnodes[2 * (size_t)t], b = tnodes[2 * (size_t)t + 1];
    const int l = asInt(a.w), r = asInt(b.w);
    const int fl = l < n ? l : (l - n < used ? newId[l - n] : -1), fr = r < n ? r : (r - n < used ? newId[r - n] : -1);
    if (fl < 0 || fr < 0 || fl >= f || fr >= f) atomicOr(status, 4);
    a.w = asFloat(fl); b.w = asFloat(fr);
    nodes[2 * (size_t)f] = a; nodes[2 * (size_t)f + 1] = b;
}

__global__ void k_hp_to_bvh2(const Vec4* __restrict__ hn, int n, Vec4* out) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    const float E = 3e37f;
    if (n == 1) {
        if (j == 0) {
            const Vec4 a = hn[0], b = hn[1];
            out[0] = v4(a.x, b.x, a.y, b.y);
            out[1] = v4(E, E, E, E);
            out[2] = v4(a.z, b.z, E, E);
            out[3] = v4(a.w, asFloat(-1), asFloat(1), asFloat(0));
        }
        return;
    }
    if (j >= n - 1) return;
    const int id = n + j;
    const int ch[2] = { asInt(hn[2 * id].w), asInt(hn[2 * id + 1].w) };
    Vec4 cl[2], chh[2];
    int c[2], k[2];
    for (int q = 0; q < 2; q++) {
        const int x = ch[q];
        if (x < 0 || x >= 2 * n - 1) {   // only an inconsistent build (reported by the builder's status) gets here
            cl[q] = v4(E, E, E, asFloat(-1)); chh[q] = v4(E, E, E, asFloat(-1)); c[q] = -1; k[q] = 0;
            continue;
        }
        cl[q] = hn[2 * x]; chh[q] = hn[2 * x + 1];
        if (x < n) { c[q] = asInt(cl[q].w); k[q] = 1; } else { c[q] = x - n; k[q] = 0; }
    }
    out[4 * j] = v4(cl[0].x, chh[0].x, cl[0].y, chh[0].y);
    out[4 * j + 1] = v4(cl[1].x, chh[1].x, cl[1].y, chh[1].y);
    out[4 * j + 2] = v4(cl[0].z, chh[0].z, cl[1].z, chh[1].z);
    out[4 * j + 3] = v4(asFloat(c[0]), asFloat(c[1]), asFloat(k[0]), asFloat(k[1]));
}

#define HCHK(x)                                                                                                        \
    do {                                                                                                               \
        hipError_t e_ = (x);                                                                                           \
        if (e_ != hipSuccess) { err = std::string(#x) + ": " + hipGetErrorString(e_); return false; }                  \
    } while (0)

}  // namespace

void HplocBuffers::sizes(int cap, bool withNodes, size_t bytes[18], int& tc) {
    const size_t n = (size_t)cap;
    // temporary node slots: the n - 1 internal nodes plus at most one partly used chunk per climbing wave
    tc = (cap - 1) + kChunk * ((cap + 62) / 64 + 1);
    const int tiles = (cap + kTile - 1) / kTile, rtiles = (tc + kTile - 1) / kTile;
    const size_t T = (size_t)tc;
    const size_t b[18] = { withNodes ? 2 * (2 * n - 1) * sizeof(Vec4) : 0,   // nodes
                           n * 8, n * 8,                                      // keysA, keysB
                           n * 4, n * 4, 4,                                   // clusters, bounds, counter
                           (size_t)256 * (std::max(tiles, rtiles) + 1) * 4,   // counts: digit x tile, then 256 totals
                           n * 2 * sizeof(Vec4),                              // cboxes
                           T * 2 * sizeof(Vec4), T * 4, T * 4, 8,             // tnodes, theight, newId, status
                           T * 8, T * 8,                                      // rkA, rkB
                           n * 4, n * 4,                                      // rangeL, rangeR
                           24, 0 };                                           // cbox
    for (int k = 0; k < 18; k++) bytes[k] = b[k];
}

HplocBuilder::~HplocBuilder() { release(); }

void HplocBuilder::release() {
    if (!b_.cap) return;
    for (void* p : { (void*)b_.nodes, (void*)b_.keysA, (void*)b_.keysB, (void*)b_.clusters, (void*)b_.bounds, (void*)b_.counter,
                     (void*)b_.counts, (void*)b_.cboxes, (void*)b_.tnodes, (void*)b_.theight, (void*)b_.newId, (void*)b_.status,
                     (void*)b_.rkA, (void*)b_.rkB, (void*)b_.rangeL, (void*)b_.rangeR, (void*)b_.cbox })
        if (p) hipFree(p);
    b_ = HplocBuffers();
}

bool HplocBuilder::reserve(int n, std::string& err) {
    if (n <= b_.cap) return true;
    if (n < 1 || n > (1 << 30)) { err = "H-PLOC: primitive count out of range"; return false; }
    release();
    size_t bytes[18];
    int tc = 0;
    HplocBuffers::sizes(n, true, bytes, tc);
    void* p[17] = {};
    for (int k = 0; k < 17; k++) {
        if (hipMalloc(&p[k], std::max<size_t>(bytes[k], 16)) != hipSuccess) {
            for (int j = 0; j < k; j++) hipFree(p[j]);
            err = "H-PLOC: out of device memory";
            return false;
        }
    }
    HplocBuffers B;
    B.nodes = (Vec4*)p[0]; B.keysA = (uint64_t*)p[1]; B.keysB = (uint64_t*)p[2]; B.clusters = (int*)p[3]; B.bounds = (int*)p[4];
    B.counter = (int*)p[5]; B.counts = (int*)p[6]; B.cboxes = (Vec4*)p[7]; B.tnodes = (Vec4*)p[8]; B.theight = (int*)p[9];
    B.newId = (int*)p[10]; B.status = (int*)p[11]; B.rkA = (uint64_t*)p[12]; B.rkB = (uint64_t*)p[13]; B.rangeL = (int*)p[14];
    B.rangeR = (int*)p[15]; B.cbox = (uint32_t*)p[16];
    B.tcap = tc; B.cap = n;
    b_ = B;
    return true;
}

bool HplocBuilder::build(const Vec4* lo, const Vec4* hi, int n, void* stream, int& root, std::string& err, int morton) {
    if (n < 1) { err = "H-PLOC: no primitives"; return false; }
    if (!lo || !hi) { err = "H-PLOC: null box arrays"; return false; }
    if (!reserve(n, err)) return false;
    return hplocEnqueue(b_, lo, hi, n, stream, morton, root, err);
}

bool hplocEnqueue(const HplocBuffers& B, const Vec4* lo, const Vec4* hi, int n, void* stream, int morton, int& root,
                  std::string& err) {
    if (n < 1 || n > B.cap) { err = "H-PLOC: primitive count outside the scratch capacity"; return false; }
    if (!validMortonMode(morton)) { err = "H-PLOC: Morton mode must be 0 (adaptive) or 30"; return false; }
    hipStream_t st = (hipStream_t)stream;
    const int g = (n + 255) / 256;
    k_hp_init<<<(B.tcap + 255) / 256, 256, 0, st>>>(B.cbox, B.counter, B.status, B.theight, B.tcap);
    k_hp_bounds<<<std::min(g, 512), 256, 0, st>>>(lo, hi, n, B.cbox);
    int ib = 1;
    while ((1ll << ib) < n) ib++;
    const int mb = morton == kMorton30 ? 10 : ((64 - ib) / 3 < 21 ? (64 - ib) / 3 : 21);
    k_hp_keys<<<g, 256, 0, st>>>(lo, hi, n, B.cbox, B.keysA, mb, ib);
    // The keys start in primitive order and their low ib bits are the primitive index, so a stable sort on the Morton
    // bits alone leaves equal codes in index order: the order of the full keys, with ib bits fewer to sort.
    const int passes = (3 * mb + 7) / 8;
    const uint64_t* sorted = B.keysA;
    if (n <= kRankSort) {
        k_rank_sort<<<(n + kSortBlock - 1) / kSortBlock, kSortBlock, 0, st>>>(B.keysA, B.keysB, n);
        sorted = B.keysB;
    } else if (n <= kSmallSort) {
        k_rs_single<<<1, kSortBlock, 0, st>>>(B.keysA, B.keysB, n, ib, passes);
        sorted = (passes & 1) ? B.keysB : B.keysA;
    } else {
        const int tiles = (n + kTile - 1) / kTile;
        uint64_t *src = B.keysA, *dst = B.keysB;
        for (int p = 0; p < passes; p++) {
            k_rs_count<<<tiles, kSortBlock, 0, st>>>(src, n, ib + 8 * p, B.counts, tiles);
            k_rs_rowscan<<<256, kSortBlock, 0, st>>>(B.counts, tiles, B.counts + 256 * (size_t)tiles);
            k_rs_scatter<<<tiles, kSortBlock, 0, st>>>(src, dst, n, ib + 8 * p, B.counts, tiles, B.counts + 256 * (size_t)tiles);
            std::swap(src, dst);
        }
        sorted = src;
    }
    k_hp_leaves_ranges<<<g, 256, 0, st>>>(lo, hi, sorted, n, ib, B.nodes, B.clusters, B.cboxes, B.bounds, B.rangeL, B.rangeR);
    if (n > 1) {
        k_hploc<<<(n - 1 + 63) / 64, 64, 0, st>>>(B.tnodes, B.theight, B.tcap, sorted, B.rangeL, B.rangeR, B.clusters, B.cboxes,
                                                  B.bounds, B.counter, n);
        // internal nodes in height order: 8-bit passes over the heights of every temporary slot (holes last); the pass
        // over the high byte runs only when a height reaches 255 (k_rn_keys sets status[1])
        // (few slots: the keys are unique, so one ranking launch orders them)
        const int T = B.tcap, rt = (T + kTile - 1) / kTile;
        k_rn_keys<<<(T + 255) / 256, 256, 0, st>>>(B.theight, T, B.rkA, B.status + 1);
        if (T <= kRankSort) {
            k_rank_sort<<<(T + kSortBlock - 1) / kSortBlock, kSortBlock, 0, st>>>(B.rkA, B.rkB, T);
            k_rn_map<<<(T + 255) / 256, 256, 0, st>>>(B.rkB, B.rkB, B.status + 1, T, n, B.newId, B.status);
        } else {
            for (int p = 0; p < 2; p++) {
                uint64_t *src = p ? B.rkB : B.rkA, *dst = p ? B.rkA : B.rkB;
                const int* gate = p ? B.status + 1 : nullptr;
                k_rs_count<<<rt, kSortBlock, 0, st>>>(src, T, 32 + 8 * p, B.counts, rt, gate);
                k_rs_rowscan<<<256, kSortBlock, 0, st>>>(B.counts, rt, B.counts + 256 * (size_t)rt, gate);
                k_rs_scatter<<<rt, kSortBlock, 0, st>>>(src, dst, T, 32 + 8 * p, B.counts, rt, B.counts + 256 * (size_t)rt, gate);
            }
            k_rn_map<<<(T + 255) / 256, 256, 0, st>>>(B.rkB, B.rkA, B.status + 1, T, n, B.newId, B.status);
        }
        k_rn_write<<<(T + 255) / 256, 256, 0, st>>>(B.tnodes, B.theight, B.newId, T, n, B.nodes, B.status);
    }
    HCHK(hipGetLastError());
    root = n > 1 ? 2 * n - 2 : 0;
    return true;
}

// ------------------------------------------------------------------------------------------------ allocation-free build
// The scratch of hplocBuildStatic: one fixed device array per GPU (the module's copy on each device), carved into the
// build's buffers on every call. Host code never names the array itself: it uses the per-device address that
// hipGetSymbolAddress returns for it.
namespace {
constexpr size_t kStaticScratchBytes = (size_t)16 << 20;   // holds the buffers of kHplocStaticMax primitives
static __device__ __align__(256) unsigned char scratch[kStaticScratchBytes];
constexpr int kMaxDevices = 64;
std::mutex g_scratchMu;
unsigned char* g_scratchPtr[kMaxDevices] = {};
}  // namespace

bool hplocBuildStatic(const Vec4* lo, const Vec4* hi, int n, int morton, Vec4* nodes, void* stream, int& root,
                      std::string& err) {
    if (n < 1 || n > kHplocStaticMax) {
        err = "H-PLOC (no allocation): " + std::to_string(n) + " primitives; the fixed scratch holds 1 to " +
              std::to_string(kHplocStaticMax);
        return false;
    }
    if (!lo || !hi || !nodes) { err = "H-PLOC (no allocation): null array"; return false; }
    int dev = 0;
    HCHK(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) { err = "H-PLOC (no allocation): device index out of range"; return false; }
    unsigned char* base = nullptr;
    {
        std::lock_guard<std::mutex> lock(g_scratchMu);
        if (!g_scratchPtr[dev]) {   // once per device, with that device current
            void* p = nullptr;
            const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
            if (e != hipSuccess || !p) {
                err = std::string("H-PLOC (no allocation): scratch address: ") + hipGetErrorString(e);
                return false;
            }
            g_scratchPtr[dev] = (unsigned char*)p;
        }
        base = g_scratchPtr[dev];
    }
    size_t bytes[18];
    int tc = 0;
    HplocBuffers::sizes(n, false, bytes, tc);
    size_t off[18], at = 0;
    for (int k = 0; k < 18; k++) { off[k] = at; at += (bytes[k] + 255) & ~(size_t)255; }
    if (at > kStaticScratchBytes) { err = "H-PLOC (no allocation): scratch too small for this build"; return false; }
    HplocBuffers B;
    B.cap = n; B.tcap = tc;
    B.nodes = nodes;
    B.keysA = (uint64_t*)(base + off[1]); B.keysB = (uint64_t*)(base + off[2]);
    B.clusters = (int*)(base + off[3]); B.bounds = (int*)(base + off[4]); B.counter = (int*)(base + off[5]);
    B.counts = (int*)(base + off[6]); B.cboxes = (Vec4*)(base + off[7]); B.tnodes = (Vec4*)(base + off[8]);
    B.theight = (int*)(base + off[9]); B.newId = (int*)(base + off[10]); B.status = (int*)(base + off[11]);
    B.rkA = (uint64_t*)(base + off[12]); B.rkB = (uint64_t*)(base + off[13]);
    B.rangeL = (int*)(base + off[14]); B.rangeR = (int*)(base + off[15]); B.cbox = (uint32_t*)(base + off[16]);
    return hplocEnqueue(B, lo, hi, n, stream, morton, root, err);
}

bool hplocToBVH2(const Vec4* hnodes, int n, Vec4* out, void* stream, std::string& err) {
    if (n < 1 || !hnodes || !out) { err = "hplocToBVH2: bad arguments"; return false; }
    const int m = n > 1 ? n - 1 : 1;
    k_hp_to_bvh2<<<(m + 127) / 128, 128, 0, (hipStream_t)stream>>>(hnodes, n, out);
    HCHK(hipGetLastError());
    return true;
}

bool hplocBuildHost(const Vec4* lo, const Vec4* hi, int n, Vec4* nodesOut, int& root, double& gpuMs, std::string& err, int morton) {
    if (n < 1) { err = "H-PLOC: no primitives"; return false; }
    Vec4 *dLo = nullptr, *dHi = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    HplocBuilder B;
    bool ok = false;
    do {
        if (hipMalloc(&dLo, (size_t)n * 16) != hipSuccess || hipMalloc(&dHi, (size_t)n * 16) != hipSuccess) { err = "H-PLOC: out of device memory"; break; }
        if (hipMemcpy(dLo, lo, (size_t)n * 16, hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(dHi, hi, (size_t)n * 16, hipMemcpyHostToDevice) != hipSuccess) { err = "H-PLOC: upload failed"; break; }
        if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) { err = "H-PLOC: event creation failed"; break; }
        if (!B.build(dLo, dHi, n, nullptr, root, err, morton)) break;   // first build: loads the kernels (not timed)
        hipEventRecord(e0, nullptr);
        if (!B.build(dLo, dHi, n, nullptr, root, err, morton)) break;
        hipEventRecord(e1, nullptr);
        if (hipEventSynchronize(e1) != hipSuccess) { err = "H-PLOC: build failed"; break; }
        float ms = 0;
        hipEventElapsedTime(&ms, e0, e1);
        gpuMs = ms;
        int status = 0;
        if (hipMemcpy(&status, B.status(), 4, hipMemcpyDeviceToHost) != hipSuccess) { err = "H-PLOC: download failed"; break; }
        if (status) { err = "H-PLOC: inconsistent build (status " + std::to_string(status) + ")"; break; }
        if (hipMemcpy(nodesOut, B.nodes(), (size_t)2 * (2 * (size_t)n - 1) * 16, hipMemcpyDeviceToHost) != hipSuccess) {
            err = "H-PLOC: download failed"; break;
        }
        ok = true;
    } while (false);
    if (e0) hipEventDestroy(e0);
    if (e1) hipEventDestroy(e1);
    if (dLo) hipFree(dLo);
    if (dHi) hipFree(dHi);
    return ok;
}

}  // namespace cwp
