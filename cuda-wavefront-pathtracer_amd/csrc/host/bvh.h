// host/bvh.h - host-side BVH construction and the GPU node/triangle layouts.
//
//  * buildSAH: recursive binned-SAH BVH2 over triangle boxes (reference: src/accel/bvh_builder.cpp:5-68), with the
//    bin count as a parameter (the reference fixes 16).
//  * layoutBVH2 / layoutBVH4: 64 B "children boxes in the parent" nodes and 128 B 4-wide nodes
//    (reference: src/accel/bvh_layouts.cpp:6-63).
//  * collapseCWBVH: SAH-optimal collapse into the 80 B compressed 8-wide layout (reference: src/accel/bvh_tools.cpp:75-168).
//  * packWoop / packVerts: triangle data in BVH leaf order (reference: src/accel/tri_pack.cpp:6-38); the vertex pack
//    keeps the three corners (not edges) so the watertight test sees bit-identical shared vertices.
#pragma once
#include "host/mesh.h"
#include <string>
#include <vector>

namespace cwp {

constexpr float kEmptyBox = 3e37f;   // coordinate of an empty child slot in BVH2/BVH4 nodes

struct BinNode {
    AABB box;
    int left = -1, right = -1, first = 0, count = 0;
    bool leaf() const { return count > 0; }
};

// Binary tree; nodes[0] is the root, ids are allocated in pre-order; leaves own prim[first, first + count).
struct BinaryBVH {
    std::vector<BinNode> nodes;
    std::vector<int> prim;
    int maxDepth = 0;
    int leafMax = 4;
};

void buildSAH(const Mesh& m, int bins, int leafMax, BinaryBVH& out);
// Binary tree from H-PLOC nodes (2 float4 per node, 2n - 1 nodes, children before parents): a bottom-up SAH choice
// per node between a leaf of all its primitives (cost A * count, only when count <= leafMax) and an internal node
// (A + cost of the children), then a pre-order tree (reference: treeFromHploc, src/accel/bvh_tools.cpp:12-38).
void treeFromHploc(const std::vector<Vec4>& hn, int n, int root, int leafMax, BinaryBVH& out);
// SAH cost normalised by the root area: sum over internal nodes A cNode + leaves A count cPrim
double sahCost(const BinaryBVH& b, double cNode, double cPrim);

// 4 float4 per node; root is node 0. Meta float4 = (c0, c1, cnt0, cnt1) as int bits: cnt > 0 leaf (c = first triangle
// position), cnt == 0 internal (c = node index, -1 = empty). A leaf root is wrapped in an internal node.
void layoutBVH2(const BinaryBVH& b, std::vector<Vec4>& nodes);
// 8 float4 per node: lo.x/hi.x/lo.y/hi.y/lo.z/hi.z for the 4 children (one child per lane), child int4, count int4.
void layoutBVH4(const BinaryBVH& b, std::vector<Vec4>& nodes);

// Compressed 8-wide BVH: 5 float4 per node, root = node 0, children of a node stored contiguously.
//  f0 = (origin.xyz, ew)  ew = (ex+127) | (ey+127) << 8 | (ez+127) << 16 | imask << 24
//  f1 = (childBase, triBase, meta[0..3], meta[4..7])   meta: 0 empty, 0x40 internal, 0x80 | (cnt-1) << 5 | offset leaf
//  f2..f4 = quantised child planes qlo.x, qlo.y, qlo.z, qhi.x, qhi.y, qhi.z (8 bytes each, two words per plane)
struct WideBVH {
    std::vector<Vec4> nodes;
    std::vector<int> prim;     // triangle order referenced by the leaves
    int maxLeafTris = 0;
    int depth = 0;             // levels of internal nodes (root alone = 1)
    double sah = 0;
    int pMax = 3; float cNode = 1.f, cPrim = 0.3f;
};
void collapseCWBVH(const BinaryBVH& b, WideBVH& out);

// depth of the binary layout's internal-node tree (the traversal stack needs at most depth - 1 entries)
int bvh2Depth(const std::vector<Vec4>& nodes);

// Woop unit-triangle rows: 3 float4 per triangle = rows of inverse([e1 e2 n]) with the translation in w (double
// precision inverse, reference: src/accel/tri_pack.cpp:18-38).
void packWoop(const Mesh& m, const std::vector<int>& order, std::vector<Vec4>& rows);
// Triangle corners: 3 float4 per triangle = (v0, original triangle id bits), (v1, material id bits), (v2, 0); and the
// per-corner normals (3 float4) when the mesh has them.
void packVerts(const Mesh& m, const std::vector<int>& order, std::vector<Vec4>& verts, std::vector<Vec4>& normals);

// Appends a BLAS layout to a shared node array, offsetting child/triangle indices; returns the BLAS root node.
int appendBVH2(std::vector<Vec4>& all, const std::vector<Vec4>& blas, int triBase);
int appendBVH4(std::vector<Vec4>& all, const std::vector<Vec4>& blas, int triBase);
int appendCWBVH(std::vector<Vec4>& all, const std::vector<Vec4>& blas, int triBase);

// Device encoding of compressed BVH8 nodes: each 80-byte node padded to 128 bytes (8 float4, one cache line), with
// the leaf-slot mask (slots whose metadata has the leaf bit) in the top byte of the child base, which replaces the
// per-slot metadata tests of the traversal. Child bases must fit 24 bits (16.7 M nodes).
bool deviceBVH8(const std::vector<Vec4>& nodes80, std::vector<Vec4>& dev, std::string& err);

}  // namespace cwp
