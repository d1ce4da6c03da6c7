// host/scene_check.h - validation of a scene in the renderer's upload layout before any of it reaches the GPU.
//
// The kernels trust the node and triangle arrays completely (no bounds checks in the traversal loop), so every scene
// passes through checkScene first: array sizes agree, every child node and leaf triangle index is in range, no node
// has two parents and the root has none (so the nodes reachable from the root form a tree and every traversal ends),
// BVH8 internal-child masks agree with the slot metadata, and material ids exist. It also returns the traversal-stack
// entries the tree can need, which sizes the per-lane spill area.
#pragma once
#include "device/renderer.h"
#include <string>

namespace cwp {

// Returns false with a message when the scene is malformed; on success sets stackNeed (entries).
bool checkScene(const SceneUpload& S, int& stackNeed, std::string& err);

}  // namespace cwp
