// device/traverse_types.h - traversal enums and records shared by host code and kernels.
#pragma once
#include "core/vecmath.h"

namespace cwp {
// triangle intersection test: Moller-Trumbore, Woop unit-triangle rows, watertight (Woop-Benthin-Wald 2013)
enum TriTest : int { kTriMT = 0, kTriWoop = 1, kTriWatertight = 2 };

// float4s per node of each host layout: BVH2 64 B, BVH4 128 B, compressed BVH8 80 B (0: unknown layout)
inline int nodeFloat4s(int variant) { return variant == 2 ? 4 : (variant == 4 ? 8 : (variant == 8 ? 5 : 0)); }

// Instance record of a two-level scene (112 bytes): world-to-object rows, object-to-world rows, BLAS root node
// (absolute index in the shared node array), mesh id, sign of det(fwd). Same layout as the reference's DevInstance
// (src/accel/instance.h:7).
struct InstanceRecord { Vec4 inv[3]; Vec4 fwd[3]; int root; int mesh; float detSign; int pad; };
static_assert(sizeof(InstanceRecord) == 112, "instance record layout");

// deepest TLAS the traversal stacks are sized for (a deeper TLAS is reported as an error, never traversed)
constexpr int kTlasMaxDepth = 48;
}  // namespace cwp
