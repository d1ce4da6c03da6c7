// host/bvh.cpp - binned-SAH BVH2 builder, BVH2/BVH4 layouts, CWBVH collapse, triangle packing.
//
// The builder and the collapse reproduce the reference's decisions exactly (same bin arithmetic in float, same strict
// comparisons, the same in-place two-sided partition, the same pre-order node numbering and the same emission order),
// so a scene's node arrays and leaf order are the reference's. Written for this library; see the per-function
// citations of src/accel/bvh_builder.cpp, src/accel/bvh_layouts.cpp and src/accel/bvh_tools.cpp.
#include "host/bvh.h"
#include <algorithm>
#include <cmath>
#include <cstring>
#include <functional>

namespace cwp {

namespace {

// Two-sided in-place partition: scan from the left for an element failing pred, from the right for one passing it,
// swap, repeat. This is the permutation the reference gets from std::partition on a bidirectional range
// (src/accel/bvh_builder.cpp:57), which fixes the triangle order inside every subtree.
template <class Pred>
int* partitionTwoSided(int* first, int* last, Pred pred) {
    for (;;) {
        for (;;) {
            if (first == last) return first;
            if (pred(*first)) ++first; else break;
        }
        --last;
        for (;;) {
            if (first == last) return first;
            if (!pred(*last)) --last; else break;
        }
        std::swap(*first, *last);
        ++first;
    }
}

struct SahBuild {
    std::vector<AABB> tb;
    std::vector<Vec3> cent;
    std::vector<int>& idx;
    std::vector<BinNode>& nodes;
    int bins, leafMax, maxDepth = 0;
    SahBuild(std::vector<int>& i, std::vector<BinNode>& n, int b, int l) : idx(i), nodes(n), bins(b), leafMax(l) {}

    int binOf(float c, float lo, float scale) const {
        int b = (int)((c - lo) * scale);
        return b < 0 ? 0 : (b > bins - 1 ? bins - 1 : b);
    }

    // src/accel/bvh_builder.cpp:14-68: leaf at <= leafMax primitives or depth 62; per axis with a non-degenerate
    // centroid extent, bin centroids, sweep right-to-left for suffix areas/counts and left-to-right for the cost
    // A_L n_L + A_R n_R (strict < keeps the first minimum); no valid split -> median split of the current order.
    int rec(int begin, int end, int depth) {
        int id = (int)nodes.size();
        nodes.push_back(BinNode());
        if (depth > maxDepth) maxDepth = depth;
        AABB box, cb;
        int n = end - begin;
        for (int i = begin; i < end; i++) { box.grow(tb[idx[i]]); cb.grow(cent[idx[i]]); }
        nodes[id].box = box;
        if (n <= leafMax || depth >= 62) { nodes[id].first = begin; nodes[id].count = n; return id; }
        float bestCost = FLT_MAX; int bestAxis = -1, bestBin = -1;
        Vec3 ext = cb.extent();
        std::vector<AABB> bb(bins); std::vector<int> bc(bins);
        std::vector<float> rA(bins); std::vector<int> rC(bins);
        for (int axis = 0; axis < 3; axis++) {
            float lo = comp(cb.lo, axis), e = comp(ext, axis);
            if (e <= 1e-12f) continue;
            float scale = (float)bins / e;
            std::fill(bb.begin(), bb.end(), AABB()); std::fill(bc.begin(), bc.end(), 0);
            for (int i = begin; i < end; i++) {
                int b = binOf(comp(cent[idx[i]], axis), lo, scale);
                bb[b].grow(tb[idx[i]]); bc[b]++;
            }
            AABB acc; int cnt = 0;
            for (int b = bins - 1; b > 0; b--) { acc.grow(bb[b]); cnt += bc[b]; rA[b] = acc.area(); rC[b] = cnt; }
            AABB lacc; int lcnt = 0;
            for (int b = 0; b < bins - 1; b++) {
                lacc.grow(bb[b]); lcnt += bc[b];
                if (lcnt == 0 || rC[b + 1] == 0) continue;
                float cost = lacc.area() * (float)lcnt + rA[b + 1] * (float)rC[b + 1];
                if (cost < bestCost) { bestCost = cost; bestAxis = axis; bestBin = b; }
            }
        }
        int mid;
        if (bestAxis < 0) {
            mid = (begin + end) / 2;   // every centroid coincides (or is degenerate on all axes): median of the order
        } else {
            float lo = comp(cb.lo, bestAxis), e = comp(ext, bestAxis);
            float scale = (float)bins / e;
            int* p = partitionTwoSided(idx.data() + begin, idx.data() + end,
                                       [&](int t) { return binOf(comp(cent[t], bestAxis), lo, scale) <= bestBin; });
            mid = (int)(p - idx.data());
            if (mid == begin || mid == end) mid = (begin + end) / 2;
        }
        int l = rec(begin, mid, depth + 1);
        int r = rec(mid, end, depth + 1);
        nodes[id].left = l; nodes[id].right = r;
        return id;
    }
};

}  // namespace

void buildSAH(const Mesh& m, int bins, int leafMax, BinaryBVH& out) {
    size_t n = m.f.size();
    out.nodes.clear(); out.prim.resize(n); out.leafMax = leafMax;
    SahBuild S(out.prim, out.nodes, bins, leafMax);
    S.tb.resize(n); S.cent.resize(n);
    for (size_t i = 0; i < n; i++) { S.tb[i] = triBox(m, i); S.cent[i] = (S.tb[i].lo + S.tb[i].hi) * 0.5f; out.prim[i] = (int)i; }
    out.nodes.reserve(2 * n / (leafMax > 0 ? leafMax : 1) + 16);
    S.rec(0, (int)n, 0);
    out.maxDepth = S.maxDepth;
}

void treeFromHploc(const std::vector<Vec4>& hn, int n, int root, int leafMax, BinaryBVH& out) {
    const int tot = 2 * n - 1;
    std::vector<int> lc(tot), rc(tot), cnt(tot);
    std::vector<double> cost(tot);
    std::vector<char> asLeaf(tot);
    std::vector<AABB> box(tot);
    for (int h = 0; h < tot; h++) {   // ids increase from children to parents: one forward sweep is bottom-up
        box[h].lo = xyz(hn[2 * (size_t)h]); box[h].hi = xyz(hn[2 * (size_t)h + 1]);
        lc[h] = asInt(hn[2 * (size_t)h].w); rc[h] = asInt(hn[2 * (size_t)h + 1].w);
        const double A = box[h].area();
        if (h < n) { cnt[h] = 1; cost[h] = A; asLeaf[h] = 1; continue; }
        cnt[h] = cnt[lc[h]] + cnt[rc[h]];
        const double inner = A + cost[lc[h]] + cost[rc[h]], leaf = cnt[h] <= leafMax ? A * cnt[h] : 1e300;
        asLeaf[h] = leaf <= inner;
        cost[h] = std::min(leaf, inner);
    }
    out.nodes.clear(); out.prim.clear(); out.maxDepth = 0; out.leafMax = leafMax;
    out.nodes.reserve(2 * (size_t)n); out.prim.reserve(n);
    struct Item { int h, parent, side, depth; };
    std::vector<Item> st = { { root, -1, 0, 0 } };
    std::vector<int> walk;
    while (!st.empty()) {
        const Item e = st.back(); st.pop_back();
        const int id = (int)out.nodes.size();
        BinNode bn; bn.box = box[e.h];
        out.nodes.push_back(bn);
        if (e.parent >= 0) (e.side ? out.nodes[e.parent].right : out.nodes[e.parent].left) = id;
        out.maxDepth = std::max(out.maxDepth, e.depth);
        if (asLeaf[e.h]) {   // the subtree's primitives, left to right
            const int first = (int)out.prim.size();
            walk.assign(1, e.h);
            while (!walk.empty()) {
                const int h = walk.back(); walk.pop_back();
                if (h < n) out.prim.push_back(lc[h]);
                else { walk.push_back(rc[h]); walk.push_back(lc[h]); }
            }
            out.nodes[id].first = first; out.nodes[id].count = (int)out.prim.size() - first;
        } else {
            st.push_back({ rc[e.h], id, 1, e.depth + 1 });
            st.push_back({ lc[e.h], id, 0, e.depth + 1 });
        }
    }
}

double sahCost(const BinaryBVH& b, double cNode, double cPrim) {
    double a0 = b.nodes[0].box.area();
    if (!(a0 > 0)) return 0;
    double s = 0;
    for (const BinNode& n : b.nodes) s += n.leaf() ? n.box.area() * n.count * cPrim : n.box.area() * cNode;
    return s / a0;
}

// ---------------------------------------------------------------- BVH2 / BVH4 layouts (src/accel/bvh_layouts.cpp:6-63)
namespace {
// A leaf root is wrapped in an internal node whose second child is empty.
std::vector<BinNode> withWrappedRoot(const BinaryBVH& b, int& root) {
    std::vector<BinNode> N = b.nodes; root = 0;
    if (N[0].leaf()) { BinNode r; r.box = N[0].box; r.left = 0; r.right = -1; N.push_back(r); root = (int)N.size() - 1; }
    return N;
}
int emit2(const std::vector<BinNode>& N, int bn, std::vector<Vec4>& out) {
    int id = (int)out.size() / 4;
    out.resize(out.size() + 4);
    int ch[2] = { N[bn].left, N[bn].right }, c[2], k[2];
    AABB bx[2];
    for (int j = 0; j < 2; j++) {
        if (ch[j] < 0) { c[j] = -1; k[j] = 0; bx[j].lo = bx[j].hi = v3(kEmptyBox, kEmptyBox, kEmptyBox); continue; }
        const BinNode& cn = N[ch[j]]; bx[j] = cn.box;
        if (cn.leaf()) { c[j] = cn.first; k[j] = cn.count; } else { k[j] = 0; c[j] = emit2(N, ch[j], out); }
    }
    Vec4* n = &out[(size_t)id * 4];
    n[0] = v4(bx[0].lo.x, bx[0].hi.x, bx[0].lo.y, bx[0].hi.y);
    n[1] = v4(bx[1].lo.x, bx[1].hi.x, bx[1].lo.y, bx[1].hi.y);
    n[2] = v4(bx[0].lo.z, bx[0].hi.z, bx[1].lo.z, bx[1].hi.z);
    n[3] = v4(asFloat(c[0]), asFloat(c[1]), asFloat(k[0]), asFloat(k[1]));
    return id;
}
int emit4(const std::vector<BinNode>& N, int bn, std::vector<Vec4>& out) {
    int id = (int)out.size() / 8;
    out.resize(out.size() + 8);
    std::vector<int> ch = { N[bn].left, N[bn].right };
    while ((int)ch.size() < 4) {   // open the largest-area internal child until 4 children
        int best = -1; float bestA = -1;
        for (size_t j = 0; j < ch.size(); j++)
            if (ch[j] >= 0 && !N[ch[j]].leaf()) { float a = N[ch[j]].box.area(); if (a > bestA) { bestA = a; best = (int)j; } }
        if (best < 0) break;
        int nb = ch[best]; ch[best] = N[nb].left; ch.push_back(N[nb].right);
    }
    int c[4], k[4]; AABB bx[4];
    for (int j = 0; j < 4; j++) {
        if (j >= (int)ch.size() || ch[j] < 0) { c[j] = -1; k[j] = 0; bx[j].lo = bx[j].hi = v3(kEmptyBox, kEmptyBox, kEmptyBox); continue; }
        const BinNode& cn = N[ch[j]]; bx[j] = cn.box;
        if (cn.leaf()) { c[j] = cn.first; k[j] = cn.count; } else { k[j] = 0; c[j] = emit4(N, ch[j], out); }
    }
    Vec4* n = &out[(size_t)id * 8];
    n[0] = v4(bx[0].lo.x, bx[1].lo.x, bx[2].lo.x, bx[3].lo.x);
    n[1] = v4(bx[0].hi.x, bx[1].hi.x, bx[2].hi.x, bx[3].hi.x);
    n[2] = v4(bx[0].lo.y, bx[1].lo.y, bx[2].lo.y, bx[3].lo.y);
    n[3] = v4(bx[0].hi.y, bx[1].hi.y, bx[2].hi.y, bx[3].hi.y);
    n[4] = v4(bx[0].lo.z, bx[1].lo.z, bx[2].lo.z, bx[3].lo.z);
    n[5] = v4(bx[0].hi.z, bx[1].hi.z, bx[2].hi.z, bx[3].hi.z);
    n[6] = v4(asFloat(c[0]), asFloat(c[1]), asFloat(c[2]), asFloat(c[3]));
    n[7] = v4(asFloat(k[0]), asFloat(k[1]), asFloat(k[2]), asFloat(k[3]));
    return id;
}
}  // namespace

void layoutBVH2(const BinaryBVH& b, std::vector<Vec4>& nodes) {
    int root; std::vector<BinNode> N = withWrappedRoot(b, root);
    nodes.clear(); emit2(N, root, nodes);
}
void layoutBVH4(const BinaryBVH& b, std::vector<Vec4>& nodes) {
    int root; std::vector<BinNode> N = withWrappedRoot(b, root);
    nodes.clear(); emit4(N, root, nodes);
}

int bvh2Depth(const std::vector<Vec4>& nodes) {
    if (nodes.empty()) return 0;
    int best = 0;
    std::vector<std::pair<int, int>> st = { { 0, 1 } };
    while (!st.empty()) {
        auto [id, d] = st.back(); st.pop_back();
        best = std::max(best, d);
        const Vec4& m = nodes[(size_t)id * 4 + 3];
        int c0 = asInt(m.x), c1 = asInt(m.y), k0 = asInt(m.z), k1 = asInt(m.w);
        if (k0 == 0 && c0 >= 0) st.push_back({ c0, d + 1 });
        if (k1 == 0 && c1 >= 0) st.push_back({ c1, d + 1 });
    }
    return best;
}

// ---------------------------------------------------------------- CWBVH collapse (src/accel/bvh_tools.cpp:75-168)
namespace {
enum : unsigned char { kDecLeaf = 0, kDecNode = 1, kDecPrev = 2, kDecDist = 10 };

int quantExp(float ext, float lo, float hi) {
    if (!(ext > 0.f)) return -100;
    int e = (int)ceilf(log2f(ext / 255.f));
    if (e < -100) e = -100;
    while (lo + 255.f * ldexpf(1.f, e) < hi) e++;
    return e;
}
unsigned pack4(const int* q, int b) {
    return (unsigned)q[b] | ((unsigned)q[b + 1] << 8) | ((unsigned)q[b + 2] << 16) | ((unsigned)q[b + 3] << 24);
}
int quant(float v, float o, float sc, bool up) {
    float f = (v - o) / sc;
    int i = up ? (int)ceilf(f) : (int)floorf(f);
    return i < 0 ? 0 : (i > 255 ? 255 : i);
}

struct Collapse {
    std::vector<AABB> box;
    std::vector<int> L, R, first, cnt;
    std::vector<float> C;            // C[n*8 + i]: cost of n as i wide-tree roots (i = 1..7)
    std::vector<unsigned char> D;    // decisions
    std::vector<unsigned char> K8;   // split of the 8 roots between the two children
    const std::vector<int>* prim = nullptr;

    void collect(int n, int i, std::vector<int>& out) const {
        for (;;) {
            if (i == 1) { out.push_back(n); return; }
            unsigned char d = D[(size_t)n * 8 + i];
            if (d == kDecPrev) { i--; continue; }
            int k = d - kDecDist;
            collect(L[n], k, out); collect(R[n], i - k, out);
            return;
        }
    }
    void prims(int n, std::vector<int>& out) const {
        std::vector<int> st(1, n);
        while (!st.empty()) {
            int h = st.back(); st.pop_back();
            if (L[h] < 0) { for (int t = 0; t < cnt[h]; t++) out.push_back((*prim)[first[h] + t]); }
            else { st.push_back(R[h]); st.push_back(L[h]); }
        }
    }
};
}  // namespace

void collapseCWBVH(const BinaryBVH& B, WideBVH& W) {
    Collapse w; w.prim = &B.prim;
    const int pMax = W.pMax; const float cNode = W.cNode, cPrim = W.cPrim;
    size_t N0 = B.nodes.size();
    w.box.resize(N0); w.L.resize(N0); w.R.resize(N0); w.first.resize(N0); w.cnt.resize(N0);
    for (size_t i = 0; i < N0; i++) {
        const BinNode& b = B.nodes[i];
        w.box[i] = b.box; w.L[i] = b.leaf() ? -1 : b.left; w.R[i] = b.leaf() ? -1 : b.right;
        w.first[i] = b.first; w.cnt[i] = b.leaf() ? b.count : 0;
    }
    // leaves above pMax primitives become binary subtrees with the same box (appended, so ids stay post-order-safe)
    for (size_t i = 0; i < w.box.size(); i++) {
        if (w.L[i] >= 0 || w.cnt[i] <= pMax) continue;
        int h = w.cnt[i] / 2, a = (int)w.box.size(), b = a + 1;
        AABB bx = w.box[i]; int f0 = w.first[i], c0 = w.cnt[i];
        for (int j = 0; j < 2; j++) { w.box.push_back(bx); w.L.push_back(-1); w.R.push_back(-1); }
        w.first.push_back(f0); w.cnt.push_back(h);
        w.first.push_back(f0 + h); w.cnt.push_back(c0 - h);
        w.L[i] = a; w.R[i] = b; w.cnt[i] = 0;
    }
    const size_t N = w.box.size();
    w.C.assign(N * 8, 0.f); w.D.assign(N * 8, 0); w.K8.assign(N, 1);
    std::vector<int> p(N, 0);
    const float INF = 3e38f;
    for (size_t ii = N; ii-- > 0;) {   // children have larger ids than their parents: decreasing id is a post-order
        int n = (int)ii;
        float A = w.box[n].area();
        float* c = &w.C[(size_t)n * 8];
        unsigned char* d = &w.D[(size_t)n * 8];
        if (w.L[n] < 0) {
            p[n] = w.cnt[n];
            float cl = A * (float)p[n] * cPrim;
            for (int i = 1; i < 8; i++) { c[i] = cl; d[i] = i == 1 ? kDecLeaf : kDecPrev; }
            continue;
        }
        int l = w.L[n], r = w.R[n];
        p[n] = p[l] + p[r];
        const float* cl_ = &w.C[(size_t)l * 8];
        const float* cr_ = &w.C[(size_t)r * 8];
        float dist[9]; int dk[9];
        for (int j = 2; j <= 8; j++) {
            dist[j] = INF; dk[j] = 1;
            for (int k = 1; k < j; k++) { float v = cl_[k] + cr_[j - k]; if (v < dist[j]) { dist[j] = v; dk[j] = k; } }
        }
        float cLeaf = p[n] <= pMax ? A * (float)p[n] * cPrim : INF;
        float cInt = dist[8] + A * cNode;
        if (cLeaf <= cInt) { c[1] = cLeaf; d[1] = kDecLeaf; } else { c[1] = cInt; d[1] = kDecNode; }
        w.K8[n] = (unsigned char)dk[8];
        for (int i = 2; i < 8; i++) {
            if (dist[i] < c[i - 1]) { c[i] = dist[i]; d[i] = (unsigned char)(kDecDist + dk[i]); }
            else { c[i] = c[i - 1]; d[i] = kDecPrev; }
        }
    }

    W.nodes.clear(); W.prim.clear(); W.prim.reserve(B.prim.size()); W.maxLeafTris = 0; W.depth = 0;
    const int root = 0;
    const float Ar = w.box[root].area();
    bool rootLeaf = w.L[root] < 0 || w.D[(size_t)root * 8 + 1] == kDecLeaf;
    if (rootLeaf) {
        float s = Ar > 0 ? (Ar * (float)p[root] * cPrim + Ar * cNode) / Ar : 0.f;
        W.sah = s;
    } else {
        float d8 = INF; int l = w.L[root], r = w.R[root];
        for (int k = 1; k < 8; k++) d8 = std::min(d8, w.C[(size_t)l * 8 + k] + w.C[(size_t)r * 8 + 8 - k]);
        float s = Ar > 0 ? (d8 + Ar * cNode) / Ar : 0.f;
        W.sah = s;
    }

    struct Job { int bn, out, level; bool asLeaf; };
    std::vector<Job> jobs;
    W.nodes.resize(5);
    jobs.push_back({ root, 0, 1, rootLeaf });
    std::vector<int> ch, pr;
    while (!jobs.empty()) {
        Job jb = jobs.back(); jobs.pop_back();
        W.depth = std::max(W.depth, jb.level);
        const int n = jb.bn;
        ch.clear();
        if (jb.asLeaf) ch.push_back(n);
        else { w.collect(w.L[n], w.K8[n], ch); w.collect(w.R[n], 8 - w.K8[n], ch); }
        const AABB box = w.box[n];
        Vec3 pc = (box.lo + box.hi) * 0.5f;
        // greedy slot assignment: repeatedly take the (child, free slot) pair of lowest cost
        // -dot(child centre - node centre, octant direction of the slot); slot bit set = negative axis
        int slotChild[8]; for (int s = 0; s < 8; s++) slotChild[s] = -1;
        float cost[8][8]; bool used[8] = { false, false, false, false, false, false, false, false };
        for (size_t c = 0; c < ch.size(); c++) {
            Vec3 cc = (w.box[ch[c]].lo + w.box[ch[c]].hi) * 0.5f - pc;
            for (int s = 0; s < 8; s++)
                cost[c][s] = -(cc.x * (s & 1 ? -1.f : 1.f) + cc.y * (s & 2 ? -1.f : 1.f) + cc.z * (s & 4 ? -1.f : 1.f));
        }
        for (size_t it = 0; it < ch.size(); it++) {
            float bc = INF; int bi = -1, bs = -1;
            for (size_t c = 0; c < ch.size(); c++)
                if (!used[c])
                    for (int s = 0; s < 8; s++)
                        if (slotChild[s] < 0 && cost[c][s] < bc) { bc = cost[c][s]; bi = (int)c; bs = s; }
            if (bi < 0)
                for (size_t c = 0; c < ch.size() && bi < 0; c++)
                    if (!used[c])
                        for (int s = 0; s < 8; s++)
                            if (slotChild[s] < 0) { bi = (int)c; bs = s; break; }
            used[bi] = true; slotChild[bs] = ch[bi];
        }
        Vec3 ext = box.extent();
        int ex = quantExp(ext.x, box.lo.x, box.hi.x), ey = quantExp(ext.y, box.lo.y, box.hi.y), ez = quantExp(ext.z, box.lo.z, box.hi.z);
        float sx = ldexpf(1.f, ex), sy = ldexpf(1.f, ey), sz = ldexpf(1.f, ez);
        int qlox[8], qloy[8], qloz[8], qhix[8], qhiy[8], qhiz[8], meta[8];
        int triBase = (int)W.prim.size(); unsigned imask = 0; int nInt = 0;
        bool childLeaf[8];
        for (int s = 0; s < 8; s++) {
            qlox[s] = qloy[s] = qloz[s] = qhix[s] = qhiy[s] = qhiz[s] = 0; meta[s] = 0; childLeaf[s] = false;
            int cn = slotChild[s];
            if (cn < 0) continue;
            const AABB& cb = w.box[cn];
            qlox[s] = quant(cb.lo.x, box.lo.x, sx, false); qhix[s] = quant(cb.hi.x, box.lo.x, sx, true);
            qloy[s] = quant(cb.lo.y, box.lo.y, sy, false); qhiy[s] = quant(cb.hi.y, box.lo.y, sy, true);
            qloz[s] = quant(cb.lo.z, box.lo.z, sz, false); qhiz[s] = quant(cb.hi.z, box.lo.z, sz, true);
            bool lf = jb.asLeaf || w.L[cn] < 0 || w.D[(size_t)cn * 8 + 1] == kDecLeaf;
            if (lf) {
                pr.clear(); w.prims(cn, pr);
                int off = (int)W.prim.size() - triBase, cnt = (int)pr.size();
                for (int t : pr) W.prim.push_back(t);
                meta[s] = 0x80 | ((cnt - 1) << 5) | off;
                W.maxLeafTris = std::max(W.maxLeafTris, cnt);
                childLeaf[s] = true;
            } else { meta[s] = 0x40; imask |= 1u << s; nInt++; }
        }
        int childBase = (int)W.nodes.size() / 5;
        W.nodes.resize(W.nodes.size() + 5 * (size_t)nInt);
        Vec4* o = &W.nodes[(size_t)jb.out * 5];
        unsigned ew = (unsigned)(ex + 127) | ((unsigned)(ey + 127) << 8) | ((unsigned)(ez + 127) << 16) | (imask << 24);
        o[0] = v4(box.lo.x, box.lo.y, box.lo.z, asFloatU(ew));
        o[1] = v4(asFloat(childBase), asFloat(triBase), asFloatU(pack4(meta, 0)), asFloatU(pack4(meta, 4)));
        unsigned u[12] = { pack4(qlox, 0), pack4(qlox, 4), pack4(qloy, 0), pack4(qloy, 4), pack4(qloz, 0), pack4(qloz, 4),
                           pack4(qhix, 0), pack4(qhix, 4), pack4(qhiy, 0), pack4(qhiy, 4), pack4(qhiz, 0), pack4(qhiz, 4) };
        for (int j = 0; j < 3; j++) o[2 + j] = v4(asFloatU(u[4 * j]), asFloatU(u[4 * j + 1]), asFloatU(u[4 * j + 2]), asFloatU(u[4 * j + 3]));
        int ord = 0;
        for (int s = 0; s < 8; s++)
            if (slotChild[s] >= 0 && !childLeaf[s]) { jobs.push_back({ slotChild[s], childBase + ord, jb.level + 1, false }); ord++; }
    }
}

// ---------------------------------------------------------------- triangle packing (src/accel/tri_pack.cpp:6-38)
void packWoop(const Mesh& m, const std::vector<int>& order, std::vector<Vec4>& rows) {
    size_t n = order.size();
    rows.resize(3 * n);
    for (size_t k = 0; k < n; k++) {
        const Tri& t = m.f[order[k]];
        Vec3 v0 = m.v[t.a], e1 = m.v[t.b] - v0, e2 = m.v[t.c] - v0, nn = cross(e1, e2);
        // columns e1, e2, n; inverse by the adjugate in double precision
        const double a00 = e1.x, a01 = e2.x, a02 = nn.x, a10 = e1.y, a11 = e2.y, a12 = nn.y, a20 = e1.z, a21 = e2.z, a22 = nn.z;
        double det = a00 * (a11 * a22 - a12 * a21) - a01 * (a10 * a22 - a12 * a20) + a02 * (a10 * a21 - a11 * a20);
        double inv[3][3] = { { 0, 0, 0 }, { 0, 0, 0 }, { 0, 0, 0 } };
        if (!(fabs(det) < 1e-300)) {
            double id = 1.0 / det;
            inv[0][0] = (a11 * a22 - a12 * a21) * id; inv[0][1] = (a02 * a21 - a01 * a22) * id; inv[0][2] = (a01 * a12 - a02 * a11) * id;
            inv[1][0] = (a12 * a20 - a10 * a22) * id; inv[1][1] = (a00 * a22 - a02 * a20) * id; inv[1][2] = (a02 * a10 - a00 * a12) * id;
            inv[2][0] = (a10 * a21 - a11 * a20) * id; inv[2][1] = (a01 * a20 - a00 * a21) * id; inv[2][2] = (a00 * a11 - a01 * a10) * id;
        }
        for (int r = 0; r < 3; r++) {
            double tr = -(inv[r][0] * v0.x + inv[r][1] * v0.y + inv[r][2] * v0.z);
            rows[3 * k + r] = v4((float)inv[r][0], (float)inv[r][1], (float)inv[r][2], (float)tr);
        }
    }
}

void packVerts(const Mesh& m, const std::vector<int>& order, std::vector<Vec4>& verts, std::vector<Vec4>& normals) {
    size_t n = order.size();
    verts.resize(3 * n);
    if (m.hasN()) normals.resize(3 * n); else normals.clear();
    for (size_t k = 0; k < n; k++) {
        int t = order[k]; const Tri& f = m.f[t];
        verts[3 * k + 0] = v4(m.v[f.a], asFloat(t));
        verts[3 * k + 1] = v4(m.v[f.b], asFloat(m.mat(t)));
        verts[3 * k + 2] = v4(m.v[f.c], 0.f);
        if (m.hasN()) { normals[3 * k] = v4(m.n[f.a], 0.f); normals[3 * k + 1] = v4(m.n[f.b], 0.f); normals[3 * k + 2] = v4(m.n[f.c], 0.f); }
    }
}

int appendBVH2(std::vector<Vec4>& all, const std::vector<Vec4>& blas, int triBase) {
    int base = (int)all.size() / 4;
    for (size_t i = 0; i < blas.size(); i += 4) {
        Vec4 mt = blas[i + 3];
        int c0 = asInt(mt.x), c1 = asInt(mt.y), k0 = asInt(mt.z), k1 = asInt(mt.w);
        if (k0 > 0) c0 += triBase; else if (c0 >= 0) c0 += base;
        if (k1 > 0) c1 += triBase; else if (c1 >= 0) c1 += base;
        all.push_back(blas[i]); all.push_back(blas[i + 1]); all.push_back(blas[i + 2]);
        all.push_back(v4(asFloat(c0), asFloat(c1), mt.z, mt.w));
    }
    return base;
}
int appendBVH4(std::vector<Vec4>& all, const std::vector<Vec4>& blas, int triBase) {
    const int base = (int)all.size() / 8;
    for (size_t i = 0; i < blas.size(); i += 8) {
        const Vec4 cf = blas[i + 6], kf = blas[i + 7];
        int c[4];
        for (int j = 0; j < 4; j++) {
            c[j] = asInt(comp(cf, j));
            const int k = asInt(comp(kf, j));
            if (k > 0) c[j] += triBase; else if (c[j] >= 0) c[j] += base;
        }
        for (int q = 0; q < 6; q++) all.push_back(blas[i + q]);
        all.push_back(v4(asFloat(c[0]), asFloat(c[1]), asFloat(c[2]), asFloat(c[3])));
        all.push_back(kf);
    }
    return base;
}
bool deviceBVH8(const std::vector<Vec4>& nodes80, std::vector<Vec4>& dev, std::string& err) {
    const size_t nn = nodes80.size() / 5;
    if (nodes80.size() % 5) { err = "BVH8: node array is not a whole number of 80-byte nodes"; return false; }
    if (nn > (size_t)1 << 24) { err = "BVH8: more than 2^24 nodes (the device encoding's child index has 24 bits)"; return false; }
    dev.assign(nn * 8, v4(0.f, 0.f, 0.f, 0.f));
    for (size_t j = 0; j < nn; j++) {
        const Vec4* n = &nodes80[j * 5];
        Vec4* o = &dev[j * 8];
        const uint32_t meta[2] = { asUint(n[1].z), asUint(n[1].w) };
        uint32_t leaf = 0;
        for (int s = 0; s < 8; s++) leaf |= ((meta[s >> 2] >> (8 * (s & 3) + 7)) & 1u) << s;
        const int childBase = asInt(n[1].x);
        if (childBase < 0 || childBase >= (1 << 24)) { err = "BVH8: child base out of the 24-bit range"; return false; }
        o[0] = n[0];
        o[1] = v4(asFloatU((uint32_t)childBase | (leaf << 24)), n[1].y, n[1].z, n[1].w);
        o[2] = n[2]; o[3] = n[3]; o[4] = n[4];
    }
    return true;
}

int appendCWBVH(std::vector<Vec4>& all, const std::vector<Vec4>& blas, int triBase) {
    int base = (int)all.size() / 5;
    for (size_t i = 0; i < blas.size(); i += 5) {
        Vec4 f1 = blas[i + 1];
        all.push_back(blas[i]);
        all.push_back(v4(asFloat(asInt(f1.x) + base), asFloat(asInt(f1.y) + triBase), f1.z, f1.w));
        all.push_back(blas[i + 2]); all.push_back(blas[i + 3]); all.push_back(blas[i + 4]);
    }
    return base;
}

}  // namespace cwp
