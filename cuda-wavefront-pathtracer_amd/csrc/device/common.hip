// device/common.hip - the kernels that do not depend on the BVH layout: split-kernel shading, pass counters, the
// per-pixel reduction, fills (declared in kernels.h).
#include "device/kernels.h"

namespace cwp {
namespace dev {

// Shading of the depth-`depth` queue from the hits of k_trace_paths (no traversal: every lane of a wave shades).
__global__ void __launch_bounds__(kBlock) k_shade(DevScene S, DevPaths P, const FrameParams* __restrict__ Fp, int depth, int smpBase, int K) {
    const FrameParams& F = *Fp;
    const int lane = laneId();
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
            const Vec4 hv = P.H[slot];
            Hit h; h.t = hv.x; h.u = hv.y; h.v = hv.z; h.tri = asInt(hv.w);
            const Vec3 T = pathThroughput(P, depth, slot);
            res = shadeLegacy(S, P, F, depth, i, i / K, (uint32_t)(smpBase + i % K), o, d, T, h);
        }
        misses += appendResults(P, cnt, part, outStart, depth, i, res);
    }
    waveAdd(misses, ctr(cnt, part, kCMiss));
}

// Continuation sort. The continuations partition x appended at depth `depth` occupy [start, start + n) of the staging
// rays P.stO / stD / stT and their keys P.qk (start: where partition x's depth-`depth` queue began; the same slots the
// unsorted append would use in ray set (depth + 1) & 1). The workgroups of partition x (blockIdx % kParts, so they run
// on that partition's XCD) walk its tiles of kSortTile entries. k_qcount adds each tile's bucket histogram (LDS
// atomics) into the partition's bucket sizes; k_qscatter gives bucket b the range after buckets 0..b-1, reserves each
// tile's share of it with one atomic per bucket and moves the rays there, tile entries in the order of their LDS ranks.
// Which entries share a wave afterwards depends on scheduling; the paths' results do not (every path keeps its own
// state and radiance).
__device__ __forceinline__ void sortRange(const DevPaths& P, int depth, int part, int& start, int& n) {
    const int* cnt = P.cnt + depth * kCntStride;
    start = partStart(cnt, part);
    n = ctrv(cnt + kCntStride, part, kCQueue);
}

__global__ void __launch_bounds__(kSortBlock) k_qcount(DevPaths P, int depth) {
    __shared__ int h[kBuckets];
    const int t = (int)threadIdx.x;
    const int part = (int)(blockIdx.x % kParts), lw = (int)(blockIdx.x / kParts), nw = (int)(gridDim.x / kParts);
    int start, n;
    sortRange(P, depth, part, start, n);
    if (t < kBuckets) h[t] = 0;
    __syncthreads();
    for (int tile = lw; tile * kSortTile < n; tile += nw) {
        const int k0 = tile * kSortTile + t;
#pragma unroll
        for (int j = 0; j < kSortPer; j++) {
            const int k = k0 + j * kSortBlock;
            if (k < n) atomicAdd(&h[P.qk[start + k] & (kBuckets - 1)], 1);
        }
    }
    __syncthreads();
    if (t < kBuckets && h[t]) atomicAdd(P.bcnt + depth * kSortWords + part * 2 * kBuckets + t, h[t]);
}

__global__ void __launch_bounds__(kSortBlock) k_qscatter(DevPaths P, int depth) {
    __shared__ int h[kBuckets], first[kBuckets], base[kBuckets];
    const int t = (int)threadIdx.x;
    const int part = (int)(blockIdx.x % kParts), lw = (int)(blockIdx.x / kParts), nw = (int)(gridDim.x / kParts);
    int start, n;
    sortRange(P, depth, part, start, n);
    int* size = P.bcnt + depth * kSortWords + part * 2 * kBuckets;
    int* cursor = size + kBuckets;
    const int set = (depth + 1) & 1;
    if (t < kBuckets) {
        // exclusive prefix of the bucket sizes over the first wave's 64 lanes
        const int c = size[t];
        int s = c;
#pragma unroll
        for (int o = 1; o < kBuckets; o <<= 1) {
            const int y = __shfl_up(s, o, kBuckets);
            if (t >= o) s += y;
        }
        first[t] = start + s - c;
    }
    for (int tile = lw; tile * kSortTile < n; tile += nw) {
        if (t < kBuckets) h[t] = 0;
        __syncthreads();
        const int k0 = tile * kSortTile + t;
        int key[kSortPer], rank[kSortPer];
#pragma unroll
        for (int j = 0; j < kSortPer; j++) {
            const int k = k0 + j * kSortBlock;
            key[j] = -1; rank[j] = 0;
            if (k < n) { key[j] = P.qk[start + k] & (kBuckets - 1); rank[j] = atomicAdd(&h[key[j]], 1); }
        }
        __syncthreads();
        if (t < kBuckets && h[t]) base[t] = first[t] + atomicAdd(cursor + t, h[t]);
        __syncthreads();
#pragma unroll
        for (int j = 0; j < kSortPer; j++)
            if (key[j] >= 0) {
                const int from = start + k0 + j * kSortBlock, to = base[key[j]] + rank[j];
                P.O[set][to] = P.stO[from]; P.D[set][to] = P.stD[from]; P.T[set][to] = P.stT[from];
            }
        __syncthreads();   // h and base are reused by the next tile
    }
}

// First-bounce sort. The depth-0 launch left each camera ray's continuation at the ray's own slot of the staging rays
// with its elevation bucket (kNoKey: none). Each tile of kFirstTile slots of partition x's camera rays (at the
// headline's 256 samples per pixel, one pixel: its bounce rays share an origin) counts its continuations per bucket
// in LDS, reserves their total in the partition's depth-1 queue with one atomic, and writes them there grouped by
// bucket, so that a depth-1 wave holds rays of one pixel that climb or descend alike and need similar traversals.
__global__ void __launch_bounds__(kFirstTile) k_qsort_first(DevPaths P) {
    __shared__ TileGroupLds g;
    const int t = (int)threadIdx.x;
    const int part = (int)(blockIdx.x % kParts), lw = (int)(blockIdx.x / kParts), nw = (int)(gridDim.x / kParts);
    const int* cnt0 = P.cnt;
    const int start = partStart(cnt0, part), n = ctrv(cnt0, part, kCQueue);
    int* queue = ctr(P.cnt + kCntStride, part, kCQueue);
    for (int tile = lw; tile * kFirstTile < n; tile += nw) {   // block-uniform bound: every thread meets the barriers
        const int k = tile * kFirstTile + t;
        const int key = k < n ? (int)P.qk[start + k] : kNoKey;
        const int to = tileGroup(key, queue, g);
        if (to >= 0) {
            const int from = start + k;
            P.O[1][start + to] = P.stO[from]; P.D[1][start + to] = P.stD[from]; P.T[1][start + to] = P.stT[from];
        }
    }
}

// zero the pass counters; the depth-0 queue of partition x is the x-th of kParts equal slices of the path ids
__device__ void passBegin(int* cnt, int words, long long NK) {
    for (int j = threadIdx.x; j < words; j += blockDim.x) cnt[j] = 0;
    __syncthreads();
    if (threadIdx.x < kParts) {
        const long long x = threadIdx.x;
        *ctr(cnt, (int)x, kCQueue) = (int)((x + 1) * NK / kParts - x * NK / kParts);
    }
}
__global__ void k_pass_begin(int* cnt, int words, long long NK) { passBegin(cnt, words, NK); }
// a batched pass: the path count comes from the pass's job table on the device
__global__ void k_pass_begin_jobs(int* cnt, int words, const DevJobTable* T) { passBegin(cnt, words, T->paths); }

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

__global__ void k_tri_geo(const Vec4* __restrict__ verts, int nt, Vec4* __restrict__ geo) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nt) return;
    const Vec4 a = verts[3 * t], b = verts[3 * t + 1], c = verts[3 * t + 2];
    geo[t] = v4(cross(xyz(b) - xyz(a), xyz(c) - xyz(a)), b.w);
}

__global__ void k_fill(uint32_t* p, size_t words, uint32_t v) {
    for (size_t j = blockIdx.x * (size_t)blockDim.x + threadIdx.x; j < words; j += (size_t)gridDim.x * blockDim.x) p[j] = v;
}

// k_reduce over the jobs of a batched pass: 64-pixel group g belongs to the last job whose first group is <= g; the
// job's accum[pix] += its K samples of the pass, in sample order (the same sums as k_reduce for the job alone)
__global__ void __launch_bounds__(kWave) k_reduce_jobs(const DevJobTable* __restrict__ T, const Vec4* __restrict__ L) {
    __shared__ float buf[kWave][3 * kRedChunk + 1];
    __shared__ int gbase[kMaxJobs];
    const int nJobs = T->nJobs, groups = T->groups;
    const DevJob* __restrict__ jobs = T->jobs;
    for (int j = threadIdx.x; j < nJobs; j += kWave) gbase[j] = jobs[j].grpBase;
    __syncthreads();
    const int lane = laneId();
    for (int g = blockIdx.x; g < groups; g += gridDim.x) {
        const int j = jobOf(gbase, nJobs, g);
        const DevJob& J = jobs[j];
        const int N = J.N, K = J.K, p0 = (g - gbase[j]) * kWave;
        const Vec4* Lj = L + J.pathBase;
        float r = 0.f, gg = 0.f, b = 0.f;
        for (int s0 = 0; s0 < K; s0 += kRedChunk) {
            const int cs = min(kRedChunk, K - s0);
            for (int q = 0; q < kWave / 4; q++) {
                const int p = 4 * q + (lane >> 4), s = lane & 15;
                if (p0 + p < N && s < cs) {
                    const Vec4 v = Lj[(size_t)(p0 + p) * K + s0 + s];
                    buf[p][3 * s] = v.x; buf[p][3 * s + 1] = v.y; buf[p][3 * s + 2] = v.z;
                }
            }
            __syncthreads();
            for (int s = 0; s < cs; s++) { r += buf[lane][3 * s]; gg += buf[lane][3 * s + 1]; b += buf[lane][3 * s + 2]; }
            __syncthreads();
        }
        const int pix = p0 + lane;
        if (pix < N) { Vec4 a = J.accum[pix]; a.x += r; a.y += gg; a.z += b; J.accum[pix] = a; }
    }
}


// ------------------------------------------------------------------------------------------------ rolling batches
__global__ void k_step_begin(const DevStepTable* T, int* cnt, int step) {
    int* blk = cnt + (step & 1) * kCntStride;
    const int* prev = cnt + ((step + 1) & 1) * kCntStride;
    for (int j = threadIdx.x; j < kCntStride; j += blockDim.x) blk[j] = 0;
    __syncthreads();
    if (threadIdx.x < kParts) {
        const long long x = threadIdx.x, C = T->camPaths;
        const int cam = (int)((x + 1) * C / kParts - x * C / kParts);
        *ctr(blk, (int)x, kCIn) = ctrv(prev, (int)x, kCQueue) + cam;
    }
}

// totals of step `step`: [0] rays traced, [1] shadow rays, [2] misses, [3] unoccluded, [4] continuations traced,
// [5] camera-ray misses (the layout of k_pass_end)
__global__ void k_step_end(const int* cnt, int step, unsigned long long* tot) {
    if (threadIdx.x != 0) return;
    const int* blk = cnt + (step & 1) * kCntStride;
    const int* prev = cnt + ((step + 1) & 1) * kCntStride;
    unsigned long long a = 0, b = 0, m = 0, u = 0, c = 0, m0 = 0;
    for (int x = 0; x < kParts; x++) {
        a += (unsigned)ctrv(blk, x, kCIn); b += (unsigned)ctrv(blk, x, kCShadow); m += (unsigned)ctrv(blk, x, kCMiss);
        u += (unsigned)ctrv(blk, x, kCUnocc); c += (unsigned)ctrv(prev, x, kCQueue); m0 += (unsigned)ctrv(blk, x, kCMissCam);
    }
    tot[0] += a; tot[1] += b; tot[2] += m; tot[3] += u; tot[4] += c; tot[5] += m0;
}

// the jobs done after this step: accum[pix] += the K samples of each pixel, in sample order (as k_reduce_jobs)
__global__ void __launch_bounds__(kWave) k_reduce_roll(const DevStepTable* __restrict__ T, const Vec4* __restrict__ L) {
    __shared__ float buf[kWave][3 * kRedChunk + 1];
    __shared__ int gbase[kMaxJobs];
    const int nDone = T->nDone, groups = T->doneGroups;
    for (int q = threadIdx.x; q < nDone; q += kWave) gbase[q] = T->doneGrp[q];
    __syncthreads();
    const int lane = laneId();
    for (int g = blockIdx.x; g < groups; g += gridDim.x) {
        const int q = jobOf(gbase, nDone, g);
        const DevJob& J = T->jobs[T->doneJob[q]];
        const int N = J.N, K = J.K, p0 = (g - gbase[q]) * kWave;
        const Vec4* Lj = L + J.pathBase;
        float r = 0.f, gg = 0.f, b = 0.f;
        for (int s0 = 0; s0 < K; s0 += kRedChunk) {
            const int cs = min(kRedChunk, K - s0);
            for (int w = 0; w < kWave / 4; w++) {
                const int p = 4 * w + (lane >> 4), s = lane & 15;
                if (p0 + p < N && s < cs) {
                    const Vec4 v = Lj[(size_t)(p0 + p) * K + s0 + s];
                    buf[p][3 * s] = v.x; buf[p][3 * s + 1] = v.y; buf[p][3 * s + 2] = v.z;
                }
            }
            __syncthreads();
            for (int s = 0; s < cs; s++) { r += buf[lane][3 * s]; gg += buf[lane][3 * s + 1]; b += buf[lane][3 * s + 2]; }
            __syncthreads();
        }
        const int pix = p0 + lane;
        if (pix < N) { Vec4 a = J.accum[pix]; a.x += r; a.y += gg; a.z += b; J.accum[pix] = a; }
    }
}

}  // namespace dev
}  // namespace cwp
