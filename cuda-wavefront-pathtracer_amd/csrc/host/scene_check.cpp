// host/scene_check.cpp - see scene_check.h.
#include "host/scene_check.h"
#include "host/bvh.h"
#include <algorithm>
#include <string>
#include <vector>

namespace cwp {

namespace {

bool fail(std::string& err, const std::string& m) { err = m; return false; }

// Child lists are checked for range, every node may have at most one parent and the root none, so the nodes
// reachable from the root form a tree and every traversal ends; depth = levels of internal nodes below the root.
struct Links {
    std::vector<int> first, kids;   // CSR child lists of internal children
};

// depth of the tree under each root; every node at most one parent, roots none, roots distinct
bool treeDepth(const Links& G, size_t nn, const std::vector<int>& roots, int& depth, std::string& err) {
    std::vector<unsigned char> parents(nn, 0);
    for (int c : G.kids) {
        if (parents[c]) return fail(err, "node " + std::to_string(c) + " has more than one parent");
        parents[c] = 1;
    }
    for (int r : roots) {
        if (parents[r]) return fail(err, "the root has a parent");
        parents[r] = 1;   // a second root at the same node is a shared tree: rejected as a second parent
    }
    depth = 0;
    for (int r : roots) {
        std::vector<std::pair<int, int>> st = { { r, 1 } };
        while (!st.empty()) {
            const auto [j, d] = st.back(); st.pop_back();
            depth = std::max(depth, d);
            for (int k = G.first[j]; k < G.first[j + 1]; k++) st.push_back({ G.kids[k], d + 1 });
        }
    }
    return true;
}

bool checkBVH2(const SceneUpload& S, size_t nn, size_t nt, Links& G, std::string& err) {
    G.first.assign(nn + 1, 0);
    for (size_t j = 0; j < nn; j++) {
        const Vec4* n = &S.nodes[j * 4];
        const int c[2] = { asInt(n[3].x), asInt(n[3].y) }, k[2] = { asInt(n[3].z), asInt(n[3].w) };
        const float box[2][6] = { { n[0].x, n[0].y, n[0].z, n[0].w, n[2].x, n[2].y },
                                  { n[1].x, n[1].y, n[1].z, n[1].w, n[2].z, n[2].w } };
        G.first[j] = (int)G.kids.size();
        for (int s = 0; s < 2; s++) {
            const std::string at = "BVH2 node " + std::to_string(j) + " child " + std::to_string(s);
            if (k[s] < 0) return fail(err, at + ": negative triangle count");
            if (k[s] > 0) {
                if (c[s] < 0 || (size_t)c[s] + (size_t)k[s] > nt) return fail(err, at + ": triangle range out of bounds");
                continue;
            }
            if (c[s] == -1) {   // empty slot: its box must be the empty box so that no ray enters it
                for (float v : box[s])
                    if (!(v >= kEmptyBox)) return fail(err, at + ": empty slot without the empty box");
                continue;
            }
            if (c[s] < 0 || (size_t)c[s] >= nn) return fail(err, at + ": child index out of range");
            G.kids.push_back(c[s]);
        }
    }
    G.first[nn] = (int)G.kids.size();
    return true;
}

bool checkBVH4(const SceneUpload& S, size_t nn, size_t nt, Links& G, std::string& err) {
    G.first.assign(nn + 1, 0);
    for (size_t j = 0; j < nn; j++) {
        const Vec4* n = &S.nodes[j * 8];
        G.first[j] = (int)G.kids.size();
        for (int s = 0; s < 4; s++) {
            const int c = asInt(comp(n[6], s)), k = asInt(comp(n[7], s));
            const std::string at = "BVH4 node " + std::to_string(j) + " child " + std::to_string(s);
            if (k < 0) return fail(err, at + ": negative triangle count");
            if (k > 0) {
                if (c < 0 || (size_t)c + (size_t)k > nt) return fail(err, at + ": triangle range out of bounds");
                continue;
            }
            if (c == -1) {
                for (int q = 0; q < 6; q++)
                    if (!(comp(n[q], s) >= kEmptyBox)) return fail(err, at + ": empty slot without the empty box");
                continue;
            }
            if (c < 0 || (size_t)c >= nn) return fail(err, at + ": child index out of range");
            G.kids.push_back(c);
        }
    }
    G.first[nn] = (int)G.kids.size();
    return true;
}

bool checkBVH8(const SceneUpload& S, size_t nn, size_t nt, Links& G, std::string& err) {
    G.first.assign(nn + 1, 0);
    for (size_t j = 0; j < nn; j++) {
        const Vec4* n = &S.nodes[j * 5];
        const uint32_t ew = asUint(n[0].w), imask = ew >> 24;
        const int childBase = asInt(n[1].x), triBase = asInt(n[1].y);
        const uint32_t meta[2] = { asUint(n[1].z), asUint(n[1].w) };
        const std::string at = "BVH8 node " + std::to_string(j);
        uint32_t internal = 0;
        for (int s = 0; s < 8; s++) {
            const uint32_t m = (meta[s >> 2] >> (8 * (s & 3))) & 255u;
            if (m == 0) continue;
            if (m & 0x80u) {
                const int first = triBase + (int)(m & 31u), cnt = (int)((m >> 5) & 3u) + 1;
                if (triBase < 0 || first < 0 || (size_t)first + (size_t)cnt > nt)
                    return fail(err, at + " slot " + std::to_string(s) + ": triangle range out of bounds");
                // the traversal gathers a node's leaf triangles into one 32-bit set relative to the triangle base
                if ((m & 31u) + (uint32_t)cnt > 32u)
                    return fail(err, at + " slot " + std::to_string(s) + ": triangle offset beyond the node's 32 triangles");
            } else internal |= 1u << s;
        }
        if (internal != imask) return fail(err, at + ": internal-child mask disagrees with the slot metadata");
        G.first[j] = (int)G.kids.size();
        const int ni = __builtin_popcount(internal);
        if (ni == 0) continue;
        if (childBase < 0 || (size_t)childBase + (size_t)ni > nn) return fail(err, at + ": children outside the node array");
        for (int c = childBase; c < childBase + ni; c++) G.kids.push_back(c);
    }
    G.first[nn] = (int)G.kids.size();
    return true;
}

}  // namespace

// full-integrator data: a material per id, emitter and environment tables of consistent sizes and ranges
// texture bindings name existing images and channels, every image lies inside the texel array, the corner uvs (and,
// with a normal map, the tangents) cover every triangle
static bool checkTextures(const SceneUpload& S, size_t nt, std::string& err) {
    if (S.maps.empty()) return true;
    if (!S.full) return fail(err, "textures need the full integrator");
    if (S.uvs.size() != 2 * nt) return fail(err, "textures: texture coordinates must be 2 float4 per triangle");
    bool normalMap = false;
    for (const MaterialMaps& x : S.maps) {
        for (int id : { x.baseColor, x.emission, x.normal, x.metallic, x.roughness })
            if (id < -1 || id >= (int)S.texImages.size()) return fail(err, "textures: image id out of range");
        if (x.metallicChannel < 0 || x.metallicChannel > 3 || x.roughnessChannel < 0 || x.roughnessChannel > 3)
            return fail(err, "textures: channel out of range");
        normalMap = normalMap || x.normal >= 0;
    }
    if (normalMap ? S.tans.size() != 3 * nt : !S.tans.empty()) return fail(err, "textures: tangents must be 3 float4 per triangle with a normal map");
    for (const TexImage& T : S.texImages)
        if (T.w < 1 || T.h < 1 || (uint64_t)T.first + (uint64_t)T.w * (uint64_t)T.h > S.texels.size() || (T.srgb != 0 && T.srgb != 1))
            return fail(err, "textures: image outside the texel array");
    if (S.texNearest != 0 && S.texNearest != 1) return fail(err, "textures: filter must be 0 or 1");
    return true;
}

static bool checkLights(const SceneUpload& S, size_t nt, std::string& err) {
    if (S.mats.size() < S.albedo.size()) return fail(err, "full integrator: fewer materials than material ids");
    if (S.lightMode < kLightMIS || S.lightMode > kLightBSDF) return fail(err, "full integrator: unknown light mode");
    if (S.rr && S.rrDepth < 0) return fail(err, "full integrator: negative roulette depth");
    for (const Material& m : S.mats)
        if (!(m.ior > 0.f) || !(m.roughness >= 0.f) || !(m.metallic >= 0.f && m.metallic <= 1.f) ||
            !(m.transmission >= 0.f && m.transmission <= 1.f))
            return fail(err, "full integrator: material parameter out of range");
    const size_t ne = S.emTri.size();
    if (ne) {
        if (S.triLightPdf.size() != nt) return fail(err, "full integrator: triLightPdf needs one entry per triangle");
        if (S.emCdf.size() != ne + 1 || S.emCdf[0] != 0.f || S.emCdf[ne] != 1.f)
            return fail(err, "full integrator: emitter CDF must have emitters + 1 entries from 0 to 1");
        for (size_t e = 0; e < ne; e++) {
            if (S.emTri[e] < 0 || (size_t)S.emTri[e] >= nt) return fail(err, "full integrator: emitter index out of range");
            if (S.emCdf[e + 1] < S.emCdf[e]) return fail(err, "full integrator: emitter CDF decreases");
        }
    }
    if (S.env && S.env->w > 0) {
        const EnvMap& E = *S.env;
        const size_t w = (size_t)E.w, h = (size_t)E.h;
        if (E.h < 1 || E.texels.size() != w * h || E.marg.size() != h + 1 || E.cond.size() != h * (w + 1))
            return fail(err, "full integrator: environment map tables have the wrong sizes");
    }
    return true;
}

bool checkScene(const SceneUpload& S, int& stackNeed, std::string& err) {
    const size_t per = nodeFloat4s(S.variant);
    if (!per) return fail(err, "variant must be 2, 4 or 8");
    if (S.nodes.empty() || S.nodes.size() % per) return fail(err, "node array size is not a whole number of nodes");
    std::vector<int> roots = S.inst ? S.meshRoot : std::vector<int>(1, S.root);
    if (roots.empty()) return fail(err, "no BLAS roots");
    for (int r : roots)
        if (r < 0 || (size_t)r >= S.nodes.size() / per) return fail(err, "root node out of range");
    if (S.inst && S.meshBox.size() != 2 * roots.size()) return fail(err, "mesh boxes must be 2 per mesh");
    if (S.verts.empty() || S.verts.size() % 3) return fail(err, "corner array size is not a whole number of triangles");
    const size_t nn = S.nodes.size() / per, nt = S.verts.size() / 3;
    if (nn > 0x7fffffff || nt > 0x7fffffff) return fail(err, "too many nodes or triangles");
    if (!S.woop.empty() && S.woop.size() != S.verts.size()) return fail(err, "Woop rows must be 3 per triangle");
    if (!S.normals.empty() && S.normals.size() != S.verts.size()) return fail(err, "normals must be 3 per triangle");
    if (S.albedo.empty()) return fail(err, "no materials");
    for (size_t t = 0; t < nt; t++) {
        const int m = asInt(S.verts[3 * t + 1].w);
        if (m < 0 || (size_t)m >= S.albedo.size())
            return fail(err, "triangle " + std::to_string(t) + ": material id out of range");
    }
    if (S.full && !checkLights(S, nt, err)) return false;
    if (!checkTextures(S, nt, err)) return false;
    int depth = 0;
    Links G;
    const bool ok = S.variant == 2 ? checkBVH2(S, nn, nt, G, err)
                                   : (S.variant == 4 ? checkBVH4(S, nn, nt, G, err) : checkBVH8(S, nn, nt, G, err));
    if (!ok || !treeDepth(G, nn, roots, depth, err)) return false;
    // a traversal defers at most one entry per level below the root (BVH2, BVH8 groups) or three (BVH4), plus one
    stackNeed = (S.variant == 4 ? 3 : 1) * depth + 1;
    return true;
}

}  // namespace cwp
