// device/traverse_types.h - traversal enums shared by host code and kernels.
#pragma once

namespace cwp {
// triangle intersection test: Moller-Trumbore, Woop unit-triangle rows, watertight (Woop-Benthin-Wald 2013)
enum TriTest : int { kTriMT = 0, kTriWoop = 1, kTriWatertight = 2 };
}  // namespace cwp
