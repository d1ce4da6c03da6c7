// device/hploc.h - GPU BVH2 builder by hierarchical PLOC (H-PLOC, Benthin et al. 2024) for gfx950.
//
// Same contract as the reference's HplocBuilder (src/accel/hploc.h:21-40): primitive boxes in, 2n - 1 nodes of two
// float4 out, (lo.xyz, left id bits), (hi.xyz, right id bits). Nodes [0, n) are the leaves in Morton order
// (left = primitive id, right = -1); internal nodes [n, 2n - 1) are numbered by height (longest path to a leaf) after
// the climb, so every child id is smaller than its parent's and the root is node 2n - 2 (node 0 when n == 1).
//
// Steps (all stream-ordered on the caller's stream, no host synchronisation):
//  1. centroid box: wave reductions + one order-preserving integer atomic per wave and axis;
//  2. 64-bit keys: Morton code (mb bits per axis) << ib | primitive index, ib = ceil(log2 n), mb = min(21, (64-ib)/3);
//  3. LSD radix sort of the keys, 8 bits per pass over the 3 mb + ib used bits (one workgroup for small n, else
//     count / scan / stable scatter passes); keys are unique, so the order is the only correct one;
//  4. leaves, then one kernel of single-wave workgroups: every lane climbs the implicit LBVH from its leaf (the
//     parent is picked by the key deltas at the two ends of the range; the first lane to reach a node stops, the
//     second continues with the union), and each LBVH node whose range exceeds 16 clusters (or the root) is reduced
//     by PLOC iterations (nearest neighbour by merged surface area within 8 positions; mutual pairs merge) to 16
//     clusters (1 at the root). A wave64 runs two such merges at once, one per half-wave of 32 lanes: the merge
//     window is 2 x 16 clusters, as in the reference's 32-lane merge, so the trees are the same. New nodes take
//     build-time slots from per-wave chunks of one counter (one atomic per chunk rather than per merge: a single
//     address hit by every merge of the grid serialises at the memory-side atomic unit);
//  5. renumbering: the internal nodes sorted by height (two radix passes), children before parents.
#pragma once
#include "core/vecmath.h"
#include <string>

namespace cwp {

// Device scratch of one build for up to `cap` primitives (the pointers of a HplocBuilder's allocations, or of a
// fixed static array for the allocation-free entry point).
struct HplocBuffers {
    int cap = 0;
    Vec4* nodes = nullptr;                        // 2n - 1 output nodes
    uint64_t *keysA = nullptr, *keysB = nullptr;
    int *clusters = nullptr, *bounds = nullptr, *counter = nullptr, *counts = nullptr;
    Vec4* cboxes = nullptr;                       // box of the cluster in each slot of clusters (2 float4), height in w
    Vec4* tnodes = nullptr;                       // internal nodes in build-time slots, renumbered into nodes
    int *theight = nullptr, *newId = nullptr, *status = nullptr;
    uint64_t *rkA = nullptr, *rkB = nullptr;
    int tcap = 0;
    int *rangeL = nullptr, *rangeR = nullptr;     // LBVH node ranges by split position
    uint32_t* cbox = nullptr;
    // bytes of each buffer for `cap` primitives, in the order of the fields above (nodes excluded when the caller
    // provides them); used both to allocate and to carve a static array
    static void sizes(int cap, bool withNodes, size_t bytes[18], int& tcap);
};

// Morton code modes: kMortonAdaptive = min(21, (64 - ceil(log2 n)) / 3) bits per axis (the reference's keys,
// src/accel/hploc.cu:181-183); kMorton30 = the classic 30-bit code, 10 bits per axis (the index keeps keys unique).
enum : int { kMortonAdaptive = 0, kMorton30 = 30 };
inline bool validMortonMode(int m) { return m == kMortonAdaptive || m == kMorton30; }

class HplocBuilder {
public:
    HplocBuilder() = default;
    ~HplocBuilder();
    HplocBuilder(const HplocBuilder&) = delete;
    HplocBuilder& operator=(const HplocBuilder&) = delete;

    bool reserve(int n, std::string& err);
    // lo / hi: n float4 per primitive box (device). stream: a hipStream_t (nullptr = default stream). Returns the root
    // node id through `root`; the nodes are ready when the stream reaches this point.
    bool build(const Vec4* lo, const Vec4* hi, int n, void* stream, int& root, std::string& err, int morton = kMortonAdaptive);
    const Vec4* nodes() const { return b_.nodes; }   // device, 2 float4 per node
    const uint64_t* sortedKeys() const { return b_.keysA; }
    // device status of the last build: 0 when the renumbered tree is consistent
    const int* status() const { return b_.status; }
    void release();
    int capacity() const { return b_.cap; }

private:
    HplocBuffers b_;
};

// One build on the given scratch (capacity >= n): every launch on `stream`, no allocation, no host synchronisation.
bool hplocEnqueue(const HplocBuffers& B, const Vec4* lo, const Vec4* hi, int n, void* stream, int morton, int& root,
                  std::string& err);

// The same build with no allocation: the scratch is a fixed static device array (16 MiB per GPU, held until the
// process ends) for at most kHplocStaticMax primitives; `nodes` is the caller's device buffer of 2n - 1 nodes. All
// calls on one GPU share the array: calls that overlap on a GPU (other streams or threads) race.
constexpr int kHplocStaticMax = 1 << 16;
bool hplocBuildStatic(const Vec4* lo, const Vec4* hi, int n, int morton, Vec4* nodes, void* stream, int& root,
                      std::string& err);

// BVH2 layout (4 float4 per node, reference layout of src/accel/bvh_layouts.h:12-13) of H-PLOC nodes on the device:
// internal node j <- H-PLOC node n + j, root n - 2; a leaf child becomes (primitive id, count 1). n == 1: node 0 holds
// the leaf and an empty slot (root 0). Reference: src/accel/hploc.cu:138-160.
bool hplocToBVH2(const Vec4* hnodes, int n, Vec4* out, void* stream, std::string& err);
inline int hplocBVH2Root(int n) { return n > 1 ? n - 2 : 0; }

// Host convenience: uploads the boxes, builds (once to warm up, once timed with events), downloads 2n - 1 nodes.
bool hplocBuildHost(const Vec4* lo, const Vec4* hi, int n, Vec4* nodesOut, int& root, double& gpuMs, std::string& err,
                    int morton = kMortonAdaptive);

}  // namespace cwp
