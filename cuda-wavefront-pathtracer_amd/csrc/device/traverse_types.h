// device/traverse_types.h - traversal enums shared by host code and kernels.
#pragma once

namespace cwp {
// triangle intersection test: Moller-Trumbore, Woop unit-triangle rows, watertight (Woop-Benthin-Wald 2013)
enum TriTest : int { kTriMT = 0, kTriWoop = 1, kTriWatertight = 2 };

// float4s per node of each host layout: BVH2 64 B, BVH4 128 B, compressed BVH8 80 B (0: unknown layout)
inline int nodeFloat4s(int variant) { return variant == 2 ? 4 : (variant == 4 ? 8 : (variant == 8 ? 5 : 0)); }
}  // namespace cwp
