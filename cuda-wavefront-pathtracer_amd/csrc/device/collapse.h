// device/collapse.h - H-PLOC nodes -> compressed BVH8 on the GPU (the host's treeFromHploc + collapseCWBVH, device
// resident), so that a scene built on the device (serving admission, per-frame rebuilds) never round-trips through
// the host.
//
// Input: the 2n - 1 H-PLOC nodes of HplocBuilder (two float4 each: (lo.xyz, left id bits), (hi.xyz, right id bits);
// leaves [0, n) name their primitive in `left`; every child id below its parent's; root 2n - 2). With one primitive per
// binary leaf (the CWBVH collapse forms its own leaves), the binary tree of treeFromHploc is the H-PLOC tree itself,
// so the collapse runs on it directly:
//  1. checks: child ids below the parent's, leaves naming primitives in [0, n), one parent per node;
//  2. the collapse's cost tables bottom-up: one lane per leaf climbs to the root, the second lane to reach a node
//     computes it from both children (cost of the subtree as 1..7 wide-tree roots, decisions, the split of 8 roots),
//     with device-coherent hand-over as in the H-PLOC climb;
//  3. the wide nodes top-down, one launch per level (its range read from device memory; the host checks once per 8
//     levels): each node's children (the decisions' roots), the greedy octant
//     slot assignment, leaf / internal flags;
//  4. the host's numbering: its job stack visits a node's internal children in decreasing slot order and appends
//     each node's children (and leaf primitives) when it pops the node, so child bases and triangle bases are prefix
//     sums over that order, computed from subtree totals (bottom-up) and offsets (top-down);
//  5. the 80-byte nodes (quantised child boxes, metadata) and the leaf primitive order.
// The arithmetic follows collapseCWBVH operation for operation (-ffp-contract=off on both sides), so the layout is
// bit-identical to the host's (reference: CWBVH::build, src/accel/bvh_tools.cpp:75-168).
#pragma once
#include "core/vecmath.h"
#include <string>
#include <vector>

namespace cwp {

class CwbvhCollapser {
public:
    CwbvhCollapser() = default;
    ~CwbvhCollapser();
    CwbvhCollapser(const CwbvhCollapser&) = delete;
    CwbvhCollapser& operator=(const CwbvhCollapser&) = delete;

    // hnodes: device H-PLOC nodes of n primitives; nodesOut: device, room for 5 float4 per node (at most n nodes);
    // primOut: device, n ints. Runs on `stream` and synchronises with it after the checks and once per 8 wide-tree levels. numNodes: the
    // wide nodes written; depth: wide-tree levels.
    bool run(const Vec4* hnodes, int n, Vec4* nodesOut, int* primOut, void* stream, int& numNodes, int& depth, std::string& err);
    // the root's cost as one wide root (after run): the collapse's SAH is this over the root box's area
    float rootCost() const { return rootCost_; }
    bool rootIsLeaf() const { return rootLeaf_; }
    void release();

private:
    bool reserve(int n, std::string& err);
    int cap_ = 0;
    void* mem_ = nullptr;   // one allocation holding every scratch array
    float rootCost_ = 0.f;
    bool rootLeaf_ = false;
    size_t bytes_ = 0;
};

// Host convenience: uploads the H-PLOC nodes, collapses on the device (timed with events), downloads the 80-byte
// nodes and the primitive order.
// sah: the collapse's SAH cost normalised by the root area (WideBVH::sah of the host collapse).
bool cwbvhFromHplocHost(const Vec4* hnodes, int n, std::vector<Vec4>& nodes, std::vector<int>& prim, int& depth,
                        double& gpuMs, float& sah, std::string& err);

}  // namespace cwp
