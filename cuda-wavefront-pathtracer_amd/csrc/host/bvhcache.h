// host/bvhcache.h - binary cache files of built scenes (BVH layout + packed triangles), so that a scene is built
// once and later setups (CLI runs, a serving process warming up) load it instead of rebuilding.
//
// A file holds the geometry part of a SceneUpload (nodes, Woop rows, corners, normals, leaf order, BLAS roots, mesh
// boxes and leaf ranges), the binary tree of the first mesh and the build statistics, under a key: a 64-bit FNV-1a
// hash of the format version, the layout and builder options and every mesh's vertices, triangles, material ids and
// normals. A file whose magic, key, sizes or trailing checksum do not match is ignored, and the next build rewrites
// it (written to a temporary name, then renamed, so a reader never sees a partial file).
#pragma once
#include "device/renderer.h"
#include "host/bvh.h"
#include "host/mesh.h"
#include <cstdint>
#include <string>
#include <vector>

namespace cwp {

struct BvhCacheEntry {
    BinaryBVH tree;          // binary tree of the first mesh (before the layout)
    int maxDepth = 0;        // deepest binary tree over the meshes
    double sah = 0;          // mean SAH cost of the layouts
};

uint64_t bvhCacheKey(const std::vector<const Mesh*>& parts, int variant, const std::string& builder, int leaf, int bins,
                     bool withNormals);
// DIR/<key as 16 hex digits>.cwpbvh
std::string bvhCachePath(const std::string& dir, uint64_t key);
// True if the file exists and is whole: then the geometry fields of S (variant, root, inst, nodes, woop, verts,
// normals, order, meshRoot, meshBox, meshTris) and E are replaced; other fields of S are untouched.
bool loadBvhCache(const std::string& path, uint64_t key, SceneUpload& S, BvhCacheEntry& E);
bool saveBvhCache(const std::string& path, uint64_t key, const SceneUpload& S, const BvhCacheEntry& E, std::string& err);

}  // namespace cwp
