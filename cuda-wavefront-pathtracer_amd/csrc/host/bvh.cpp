// host/bvh.cpp - host BVH construction: binned-SAH BVH2, the BVH2 / BVH4 node layouts, the compressed BVH8 collapse,
// the SAH leaf collapse of H-PLOC trees, and triangle packing.
//
// Every scene's node arrays must be the ones the reference's algorithms produce (the oracle checks them bit for bit),
// so the DECISIONS follow the reference (src/accel/bvh_builder.cpp:14-68, bvh_layouts.cpp:6-63, bvh_tools.cpp:12-168,
// tri_pack.cpp:18-38): the same binning arithmetic in float, strict comparisons that keep the first minimum, the
// permutation of a two-sided in-place partition, pre-order numbering, and the 8-wide collapse's cost recurrence, slot
// assignment and quantisation. The code is organised differently: the builder runs from an explicit task stack and
// bins all three axes in one pass over a range; the layouts first number the nodes (one pre-order walk), then fill
// every node independently; the 8-wide collapse runs in three phases (cost tables; the wide tree's shape level by
// level; the numbering of nodes and triangles in the order a depth-first job stack would visit them, then the
// quantised records), the structure the GPU collapse (device/collapse.hip) also has.
#include "host/bvh.h"
#include <algorithm>
#include <cmath>
#include <cstring>

namespace cwp {

// ------------------------------------------------------------------------------------------------ binned-SAH BVH2
namespace {

// Bin of a centroid coordinate c on an axis whose centroid range starts at lo, with scale = bins / extent.
inline int centroidBin(float c, float lo, float scale, int bins) {
    const int b = (int)((c - lo) * scale);
    return std::min(std::max(b, 0), bins - 1);
}

// Moves the entries of idx[begin, end) that satisfy keep() in front of the others, exchanging the leftmost entry that
// fails with the rightmost entry that passes until the two scans meet. This is the permutation of a bidirectional
// std::partition, which fixes the primitive order inside every subtree of the reference's builder.
template <class Keep>
int partitionRange(std::vector<int>& idx, int begin, int end, Keep keep) {
    int i = begin, j = end;   // [begin, i) kept, [j, end) rejected
    for (;;) {
        while (i < j && keep(idx[i])) i++;
        while (i < j && !keep(idx[j - 1])) j--;
        if (i >= j) return i;
        std::swap(idx[i], idx[j - 1]);
        i++; j--;
    }
}

class SahBuilder {
public:
    SahBuilder(const Mesh& m, int bins, int leafMax, BinaryBVH& out) : bins_(bins), leafMax_(leafMax), out_(out) {
        const size_t n = m.f.size();
        box_.resize(n); cen_.resize(n);
        out_.prim.resize(n);
        for (size_t t = 0; t < n; t++) { box_[t] = triBox(m, t); cen_[t] = (box_[t].lo + box_[t].hi) * 0.5f; out_.prim[t] = (int)t; }
        for (int a = 0; a < 3; a++) { binBox_[a].resize(bins); binCnt_[a].resize(bins); }
        sufArea_.resize(bins); sufCnt_.resize(bins);
    }

    void run() {
        out_.nodes.clear();
        out_.nodes.reserve(2 * out_.prim.size() / (leafMax_ > 0 ? leafMax_ : 1) + 16);
        out_.maxDepth = 0;
        // pre-order numbering: a task takes the next node id when it is popped; the right half is pushed first so
        // that the whole left subtree is numbered before it
        struct Task { int begin, end, depth, parent; bool right; };
        std::vector<Task> todo{ { 0, (int)out_.prim.size(), 0, -1, false } };
        while (!todo.empty()) {
            const Task t = todo.back();
            todo.pop_back();
            const int id = (int)out_.nodes.size();
            out_.nodes.emplace_back();
            if (t.parent >= 0) (t.right ? out_.nodes[t.parent].right : out_.nodes[t.parent].left) = id;
            out_.maxDepth = std::max(out_.maxDepth, t.depth);
            const int mid = split(id, t.begin, t.end, t.depth);
            if (mid < 0) continue;
            todo.push_back({ mid, t.end, t.depth + 1, id, true });
            todo.push_back({ t.begin, mid, t.depth + 1, id, false });
        }
    }

private:
    // Bounds node `id` over idx[begin, end) and either makes it a leaf (returns -1) or partitions the range and returns
    // the split position: binned SAH (cost A_L n_L + A_R n_R over centroid bins, first minimum over axes 0, 1, 2 and
    // bins in order); every centroid on one point: the middle of the current order.
    int split(int id, int begin, int end, int depth) {
        std::vector<int>& idx = out_.prim;
        AABB box, cb;
        for (int k = begin; k < end; k++) { box.grow(box_[idx[k]]); cb.grow(cen_[idx[k]]); }
        BinNode& nd = out_.nodes[id];
        nd.box = box;
        const int n = end - begin;
        if (n <= leafMax_ || depth >= 62) { nd.first = begin; nd.count = n; return -1; }
        const Vec3 ext = cb.extent();
        float scale[3];
        bool live[3];
        for (int a = 0; a < 3; a++) {
            live[a] = !(comp(ext, a) <= 1e-12f);   // a degenerate axis has no split (a NaN extent is not skipped)
            scale[a] = live[a] ? (float)bins_ / comp(ext, a) : 0.f;
            if (!live[a]) continue;
            std::fill(binBox_[a].begin(), binBox_[a].end(), AABB());
            std::fill(binCnt_[a].begin(), binCnt_[a].end(), 0);
        }
        // one pass over the range fills the bins of every live axis
        for (int k = begin; k < end; k++) {
            const int t = idx[k];
            for (int a = 0; a < 3; a++) {
                if (!live[a]) continue;
                const int b = centroidBin(comp(cen_[t], a), comp(cb.lo, a), scale[a], bins_);
                binBox_[a][b].grow(box_[t]);
                binCnt_[a][b]++;
            }
        }
        float best = FLT_MAX;
        int bestAxis = -1, bestBin = -1;
        for (int a = 0; a < 3; a++) {
            if (!live[a]) continue;
            // suffix areas and counts of bins b.. bins - 1, then the left sweep
            AABB r; int rc = 0;
            for (int b = bins_ - 1; b > 0; b--) { r.grow(binBox_[a][b]); rc += binCnt_[a][b]; sufArea_[b] = r.area(); sufCnt_[b] = rc; }
            AABB l; int lc = 0;
            for (int b = 0; b + 1 < bins_; b++) {
                l.grow(binBox_[a][b]); lc += binCnt_[a][b];
                if (lc == 0 || sufCnt_[b + 1] == 0) continue;
                const float cost = l.area() * (float)lc + sufArea_[b + 1] * (float)sufCnt_[b + 1];
                if (cost < best) { best = cost; bestAxis = a; bestBin = b; }
            }
        }
        int mid = (begin + end) / 2;
        if (bestAxis >= 0) {
            const float lo = comp(cb.lo, bestAxis), sc = scale[bestAxis];
            const int ax = bestAxis, bb = bestBin;
            const int p = partitionRange(idx, begin, end, [&](int t) { return centroidBin(comp(cen_[t], ax), lo, sc, bins_) <= bb; });
            if (p != begin && p != end) mid = p;
        }
        return mid;
    }

    int bins_, leafMax_;
    BinaryBVH& out_;
    std::vector<AABB> box_;
    std::vector<Vec3> cen_;
    std::vector<AABB> binBox_[3];
    std::vector<int> binCnt_[3];
    std::vector<float> sufArea_;
    std::vector<int> sufCnt_;
};

}  // namespace

void buildSAH(const Mesh& m, int bins, int leafMax, BinaryBVH& out) {
    out.leafMax = leafMax;
    SahBuilder b(m, bins, leafMax, out);
    b.run();
}

// ------------------------------------------------------------------------------------------------ H-PLOC leaf collapse
void treeFromHploc(const std::vector<Vec4>& hn, int n, int root, int leafMax, BinaryBVH& out) {
    const int tot = 2 * n - 1;
    auto child = [&](int h, int side) { return asInt(hn[2 * (size_t)h + side].w); };
    auto boxOf = [&](int h) { AABB b; b.lo = xyz(hn[2 * (size_t)h]); b.hi = xyz(hn[2 * (size_t)h + 1]); return b; };
    // 1. SAH cost of the best subtree at each node (double): leaf of all its primitives (A * count, allowed up to
    //    leafMax) or internal (A + both children); H-PLOC ids put children below parents, so ascending ids are bottom-up
    std::vector<int> count(tot);
    std::vector<double> best(tot);
    std::vector<char> leafHere(tot);
    for (int h = 0; h < tot; h++) {
        const double A = boxOf(h).area();
        if (h < n) { count[h] = 1; best[h] = A; leafHere[h] = 1; continue; }
        const int l = child(h, 0), r = child(h, 1);
        count[h] = count[l] + count[r];
        const double asInner = A + best[l] + best[r];
        const double asLeaf = count[h] <= leafMax ? A * count[h] : 1e300;
        leafHere[h] = asLeaf <= asInner;
        best[h] = std::min(asLeaf, asInner);
    }
    // 2. the primitives in left-to-right leaf order, and where each node's primitives start in it: a collapsed leaf
    //    owns a contiguous run of that order, and the pre-order emission below meets the leaves left to right
    std::vector<int> start(tot, 0);
    out.prim.clear(); out.prim.reserve(n);
    {
        std::vector<int> walk{ root };
        while (!walk.empty()) {
            const int h = walk.back(); walk.pop_back();
            start[h] = (int)out.prim.size();
            if (h < n) out.prim.push_back(child(h, 0));
            else { walk.push_back(child(h, 1)); walk.push_back(child(h, 0)); }
        }
    }
    // 3. pre-order emission down to the collapsed leaves
    out.nodes.clear(); out.nodes.reserve(2 * (size_t)n); out.maxDepth = 0; out.leafMax = leafMax;
    struct Visit { int h, parent, depth; bool right; };
    std::vector<Visit> todo{ { root, -1, 0, false } };
    while (!todo.empty()) {
        const Visit v = todo.back(); todo.pop_back();
        const int id = (int)out.nodes.size();
        out.nodes.emplace_back();
        out.nodes[id].box = boxOf(v.h);
        if (v.parent >= 0) (v.right ? out.nodes[v.parent].right : out.nodes[v.parent].left) = id;
        out.maxDepth = std::max(out.maxDepth, v.depth);
        if (leafHere[v.h]) { out.nodes[id].first = start[v.h]; out.nodes[id].count = count[v.h]; continue; }
        todo.push_back({ child(v.h, 1), id, v.depth + 1, true });
        todo.push_back({ child(v.h, 0), id, v.depth + 1, false });
    }
}

double sahCost(const BinaryBVH& b, double cNode, double cPrim) {
    double a0 = b.nodes[0].box.area();
    if (!(a0 > 0)) return 0;
    double s = 0;
    for (const BinNode& n : b.nodes) s += n.leaf() ? n.box.area() * n.count * cPrim : n.box.area() * cNode;
    return s / a0;
}

// ------------------------------------------------------------------------------------------------ BVH2 / BVH4 layouts
namespace {

const AABB kNoBox = [] { AABB b; b.lo = b.hi = v3(kEmptyBox, kEmptyBox, kEmptyBox); return b; }();

// A child slot of a layout node: an internal child (its layout id), a leaf (first triangle, count) or empty.
struct Slot { AABB box = kNoBox; int ref = -1, count = 0; };

// The binary tree with a leaf root wrapped in an internal node whose second child is empty (the layouts need an
// internal root).
struct Rooted {
    const std::vector<BinNode>& N;
    BinNode wrap;
    int root;
    explicit Rooted(const BinaryBVH& b) : N(b.nodes), root(0) {
        if (N[0].leaf()) { wrap.box = N[0].box; wrap.left = 0; wrap.right = -1; root = -2; }
    }
    const BinNode& at(int i) const { return i == -2 ? wrap : N[i]; }
};

// Children of a W-wide layout node over binary node bn: W = 2 takes the two children; W = 4 keeps opening the
// internal child of largest area (first of equal areas) until there are 4 children or only leaves remain.
std::vector<int> wideChildren(const Rooted& T, int bn, int W) {
    std::vector<int> ch{ T.at(bn).left, T.at(bn).right };
    while (W > 2 && (int)ch.size() < W) {
        int pick = -1;
        float area = -1.f;
        for (size_t j = 0; j < ch.size(); j++) {
            if (ch[j] < 0 || T.at(ch[j]).leaf()) continue;
            const float a = T.at(ch[j]).box.area();
            if (a > area) { area = a; pick = (int)j; }
        }
        if (pick < 0) break;
        const BinNode& o = T.at(ch[pick]);
        ch[pick] = o.left;
        ch.push_back(o.right);
    }
    return ch;
}

// Numbers the layout nodes in pre-order (children in slot order) and returns, per layout node, its binary node and
// its child slots.
void layoutPlan(const BinaryBVH& b, int W, std::vector<std::vector<Slot>>& slots) {
    const Rooted T(b);
    std::vector<int> order;                 // binary node of each layout id
    std::vector<std::vector<int>> kids;     // binary children of each layout id
    std::vector<int> todo{ T.root };
    while (!todo.empty()) {
        const int bn = todo.back(); todo.pop_back();
        order.push_back(bn);
        kids.push_back(wideChildren(T, bn, W));
        const std::vector<int>& ch = kids.back();
        for (size_t j = ch.size(); j-- > 0;)
            if (ch[j] >= 0 && !T.at(ch[j]).leaf()) todo.push_back(ch[j]);
    }
    // layout id of every internal binary node that heads a layout node
    std::vector<int> idOf(b.nodes.size(), -1);
    for (size_t k = 0; k < order.size(); k++) if (order[k] >= 0) idOf[order[k]] = (int)k;
    slots.assign(order.size(), std::vector<Slot>(W));
    for (size_t k = 0; k < order.size(); k++)
        for (size_t j = 0; j < kids[k].size() && (int)j < W; j++) {
            const int c = kids[k][j];
            if (c < 0) continue;
            Slot& s = slots[k][j];
            s.box = T.at(c).box;
            if (T.at(c).leaf()) { s.ref = T.at(c).first; s.count = T.at(c).count; }
            else s.ref = idOf[c];
        }
}

}  // namespace

void layoutBVH2(const BinaryBVH& b, std::vector<Vec4>& nodes) {
    std::vector<std::vector<Slot>> plan;
    layoutPlan(b, 2, plan);
    nodes.assign(plan.size() * 4, v4(0.f, 0.f, 0.f, 0.f));
    for (size_t k = 0; k < plan.size(); k++) {
        const Slot &a = plan[k][0], &c = plan[k][1];
        Vec4* o = &nodes[k * 4];
        o[0] = v4(a.box.lo.x, a.box.hi.x, a.box.lo.y, a.box.hi.y);
        o[1] = v4(c.box.lo.x, c.box.hi.x, c.box.lo.y, c.box.hi.y);
        o[2] = v4(a.box.lo.z, a.box.hi.z, c.box.lo.z, c.box.hi.z);
        o[3] = v4(asFloat(a.ref), asFloat(c.ref), asFloat(a.count), asFloat(c.count));
    }
}

void layoutBVH4(const BinaryBVH& b, std::vector<Vec4>& nodes) {
    std::vector<std::vector<Slot>> plan;
    layoutPlan(b, 4, plan);
    nodes.assign(plan.size() * 8, v4(0.f, 0.f, 0.f, 0.f));
    for (size_t k = 0; k < plan.size(); k++) {
        const std::vector<Slot>& s = plan[k];
        Vec4* o = &nodes[k * 8];
        auto row = [&](auto f) { return v4(f(s[0]), f(s[1]), f(s[2]), f(s[3])); };
        o[0] = row([](const Slot& x) { return x.box.lo.x; }); o[1] = row([](const Slot& x) { return x.box.hi.x; });
        o[2] = row([](const Slot& x) { return x.box.lo.y; }); o[3] = row([](const Slot& x) { return x.box.hi.y; });
        o[4] = row([](const Slot& x) { return x.box.lo.z; }); o[5] = row([](const Slot& x) { return x.box.hi.z; });
        o[6] = row([](const Slot& x) { return asFloat(x.ref); }); o[7] = row([](const Slot& x) { return asFloat(x.count); });
    }
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

// ------------------------------------------------------------------------------------------------ compressed BVH8
namespace {

constexpr float kCostInf = 3e38f;

// Per binary node and root count i (1..7): the cheapest way to cover the node's subtree with i wide-tree roots, and
// how: i = 1 as a leaf or as an internal wide node (whose 8 roots split k8 / 8 - k8 between the children); i >= 2 as
// the best i - 1 cover, or k roots in the left child and i - k in the right.
struct CoverTable {
    enum : unsigned char { kLeaf = 0, kNode = 1, kFewer = 2, kSplit = 16 };   // kSplit + k: k roots on the left
    std::vector<float> cost;          // [node * 8 + i]
    std::vector<unsigned char> how;   // [node * 8 + i]
    std::vector<unsigned char> k8;    // per node
    bool leaf(int n) const { return how[(size_t)n * 8 + 1] == kLeaf; }
};

// Binary tree of the collapse: the input's nodes, with leaves above pMax primitives split into halves (new nodes
// appended, so every child id stays above its parent's).
struct CollapseTree {
    std::vector<AABB> box;
    std::vector<int> left, right, first, count;   // left < 0: leaf of prim[first, first + count)
    int size() const { return (int)box.size(); }
    void add(const AABB& b, int l, int r, int f, int c) { box.push_back(b); left.push_back(l); right.push_back(r); first.push_back(f); count.push_back(c); }
};

CollapseTree collapseInput(const BinaryBVH& B, int pMax) {
    CollapseTree T;
    for (const BinNode& b : B.nodes) T.add(b.box, b.leaf() ? -1 : b.left, b.leaf() ? -1 : b.right, b.first, b.leaf() ? b.count : 0);
    for (int i = 0; i < T.size(); i++) {
        if (T.left[i] >= 0 || T.count[i] <= pMax) continue;
        const int h = T.count[i] / 2, a = T.size();
        const AABB bx = T.box[i];
        const int f0 = T.first[i], c0 = T.count[i];
        T.add(bx, -1, -1, f0, h);
        T.add(bx, -1, -1, f0 + h, c0 - h);
        T.left[i] = a; T.right[i] = a + 1; T.count[i] = 0;
    }
    return T;
}

// Phase 1: cover costs bottom-up (descending ids); also each node's primitive count.
void coverCosts(const CollapseTree& T, int pMax, float cNode, float cPrim, CoverTable& C, std::vector<int>& prims) {
    const int N = T.size();
    C.cost.assign((size_t)N * 8, 0.f); C.how.assign((size_t)N * 8, CoverTable::kFewer); C.k8.assign(N, 1);
    prims.assign(N, 0);
    for (int n = N - 1; n >= 0; n--) {
        const float A = T.box[n].area();
        float* c = &C.cost[(size_t)n * 8];
        unsigned char* how = &C.how[(size_t)n * 8];
        if (T.left[n] < 0) {
            prims[n] = T.count[n];
            const float asLeaf = A * (float)prims[n] * cPrim;
            for (int i = 1; i < 8; i++) c[i] = asLeaf;
            how[1] = CoverTable::kLeaf;
            continue;
        }
        const float* cl = &C.cost[(size_t)T.left[n] * 8];
        const float* cr = &C.cost[(size_t)T.right[n] * 8];
        prims[n] = prims[T.left[n]] + prims[T.right[n]];
        // split[j]: best k roots left + j - k right, j = 2..8 (first k of the minimum)
        float split[9];
        int splitK[9];
        for (int j = 2; j <= 8; j++) {
            split[j] = kCostInf; splitK[j] = 1;
            for (int k = 1; k < j; k++) {
                const float v = cl[k] + cr[j - k];
                if (v < split[j]) { split[j] = v; splitK[j] = k; }
            }
        }
        const float asLeaf = prims[n] <= pMax ? A * (float)prims[n] * cPrim : kCostInf;
        const float asNode = split[8] + A * cNode;
        if (asLeaf <= asNode) { c[1] = asLeaf; how[1] = CoverTable::kLeaf; }
        else { c[1] = asNode; how[1] = CoverTable::kNode; }
        C.k8[n] = (unsigned char)splitK[8];
        for (int i = 2; i < 8; i++) {
            if (split[i] < c[i - 1]) { c[i] = split[i]; how[i] = (unsigned char)(CoverTable::kSplit + splitK[i]); }
            else { c[i] = c[i - 1]; how[i] = CoverTable::kFewer; }
        }
    }
}

// the i roots that cover node n, left to right
void coverRoots(const CollapseTree& T, const CoverTable& C, int n, int i, std::vector<int>& out) {
    std::vector<std::pair<int, int>> todo{ { n, i } };   // processed last-in-first-out: push right before left
    while (!todo.empty()) {
        auto [m, k] = todo.back(); todo.pop_back();
        while (k > 1 && C.how[(size_t)m * 8 + k] == CoverTable::kFewer) k--;
        if (k == 1) { out.push_back(m); continue; }
        const int kl = C.how[(size_t)m * 8 + k] - CoverTable::kSplit;
        todo.push_back({ T.right[m], k - kl });
        todo.push_back({ T.left[m], kl });
    }
}

// A node of the wide tree: the binary node it stands for and what sits in its 8 slots.
struct WideNode {
    int bin;
    int child[8];      // binary node in the slot, -1 = empty
    bool isLeaf[8];    // the slot holds a leaf (its subtree's primitives) rather than a wide node
    int wideChild[8];  // index of the slot's wide node (internal slots)
    int level;
};

// Greedy octant slot assignment: repeatedly the (child, free slot) pair of lowest cost -dot(child centre - node
// centre, octant direction of the slot) (slot bit k set = negative axis k), children in order, first minimum.
void assignSlots(const CollapseTree& T, int parentBin, const std::vector<int>& ch, int slotOf[8]) {
    const Vec3 pc = (T.box[parentBin].lo + T.box[parentBin].hi) * 0.5f;
    const int m = (int)ch.size();
    float cost[8][8];
    for (int c = 0; c < m; c++) {
        const Vec3 d = (T.box[ch[c]].lo + T.box[ch[c]].hi) * 0.5f - pc;
        for (int s = 0; s < 8; s++)
            cost[c][s] = -(d.x * (s & 1 ? -1.f : 1.f) + d.y * (s & 2 ? -1.f : 1.f) + d.z * (s & 4 ? -1.f : 1.f));
    }
    bool placed[8] = { false, false, false, false, false, false, false, false };
    for (int s = 0; s < 8; s++) slotOf[s] = -1;
    for (int round = 0; round < m; round++) {
        float lowest = kCostInf;
        int bc = -1, bs = -1;
        for (int c = 0; c < m; c++) {
            if (placed[c]) continue;
            for (int s = 0; s < 8; s++)
                if (slotOf[s] < 0 && cost[c][s] < lowest) { lowest = cost[c][s]; bc = c; bs = s; }
        }
        if (bc < 0) {   // no finite cost left (non-finite boxes): the first unplaced child in the first free slot
            for (int c = 0; c < m && bc < 0; c++) if (!placed[c]) bc = c;
            for (int s = 0; s < 8 && bs < 0; s++) if (slotOf[s] < 0) bs = s;
        }
        placed[bc] = true;
        slotOf[bs] = bc;
    }
}

// exponent of a 2^e grid on [lo, hi] with 256 steps: the smallest e >= ceil(log2(ext / 255)) (at least -100) with
// lo + 255 2^e >= hi
int gridExponent(float ext, float lo, float hi) {
    if (!(ext > 0.f)) return -100;
    int e = std::max((int)ceilf(log2f(ext / 255.f)), -100);
    while (lo + 255.f * ldexpf(1.f, e) < hi) e++;
    return e;
}
uint32_t gridStep(float v, float org, float step, bool up) {
    const float q = (v - org) / step;
    const int i = up ? (int)ceilf(q) : (int)floorf(q);
    return (uint32_t)std::min(std::max(i, 0), 255);
}

}  // namespace

void collapseCWBVH(const BinaryBVH& B, WideBVH& W) {
    const CollapseTree T = collapseInput(B, W.pMax);
    CoverTable C;
    std::vector<int> prims;
    coverCosts(T, W.pMax, W.cNode, W.cPrim, C, prims);
    const int root = 0;
    const bool rootLeaf = T.left[root] < 0 || C.leaf(root);
    {   // SAH of the wide tree, normalised by the root area
        const float Ar = T.box[root].area();
        float total;
        if (rootLeaf) total = Ar * (float)prims[root] * W.cPrim + Ar * W.cNode;
        else {
            float d8 = kCostInf;
            for (int k = 1; k < 8; k++) d8 = std::min(d8, C.cost[(size_t)T.left[root] * 8 + k] + C.cost[(size_t)T.right[root] * 8 + 8 - k]);
            total = d8 + Ar * W.cNode;
        }
        W.sah = Ar > 0 ? total / Ar : 0.f;
    }

    // Phase 2: the wide tree's shape, level by level: each wide node's cover roots placed in octant slots
    std::vector<WideNode> wide;
    std::vector<int> ch;
    auto makeWide = [&](int bin, bool allLeaf, int level) {
        WideNode w;
        w.bin = bin; w.level = level;
        ch.clear();
        if (allLeaf) ch.push_back(bin);
        else { coverRoots(T, C, T.left[bin], C.k8[bin], ch); coverRoots(T, C, T.right[bin], 8 - C.k8[bin], ch); }
        int slotOf[8];
        assignSlots(T, bin, ch, slotOf);
        for (int s = 0; s < 8; s++) {
            w.child[s] = slotOf[s] < 0 ? -1 : ch[slotOf[s]];
            w.isLeaf[s] = w.child[s] >= 0 && (allLeaf || T.left[w.child[s]] < 0 || C.leaf(w.child[s]));
            w.wideChild[s] = -1;
        }
        wide.push_back(w);
    };
    makeWide(root, rootLeaf, 1);
    for (size_t k = 0; k < wide.size(); k++)   // breadth first: children are appended behind their level
        for (int s = 0; s < 8; s++)
            if (wide[k].child[s] >= 0 && !wide[k].isLeaf[s]) {
                const int idx = (int)wide.size();
                makeWide(wide[k].child[s], false, wide[k].level + 1);
                wide[k].wideChild[s] = idx;
            }

    // Phase 3: numbering in depth-first job order (a node's internal children get consecutive node ids in slot
    // order when the node is visited; visits continue with its last internal child first), which also fixes the
    // order in which the leaves' triangles are appended.
    std::vector<int> nodeId(wide.size(), -1), childBase(wide.size(), 0), triBase(wide.size(), 0);
    std::vector<std::vector<int>> leafTris(wide.size());
    W.prim.clear(); W.prim.reserve(B.prim.size());
    W.maxLeafTris = 0; W.depth = 0;
    {
        int nodes = 1;
        nodeId[0] = 0;
        std::vector<int> todo{ 0 }, walk;
        while (!todo.empty()) {
            const int k = todo.back(); todo.pop_back();
            const WideNode& w = wide[k];
            W.depth = std::max(W.depth, w.level);
            triBase[k] = (int)W.prim.size();
            childBase[k] = nodes;
            for (int s = 0; s < 8; s++) {
                if (w.child[s] < 0) continue;
                if (!w.isLeaf[s]) { nodeId[w.wideChild[s]] = nodes++; todo.push_back(w.wideChild[s]); continue; }
                // the leaf's primitives: the binary subtree's leaves left to right
                const int before = (int)W.prim.size();
                walk.assign(1, w.child[s]);
                while (!walk.empty()) {
                    const int h = walk.back(); walk.pop_back();
                    if (T.left[h] < 0) for (int t = 0; t < T.count[h]; t++) W.prim.push_back(B.prim[T.first[h] + t]);
                    else { walk.push_back(T.right[h]); walk.push_back(T.left[h]); }
                }
                leafTris[k].push_back((int)W.prim.size() - before);
                W.maxLeafTris = std::max(W.maxLeafTris, (int)W.prim.size() - before);
            }
        }
        W.nodes.assign((size_t)nodes * 5, v4(0.f, 0.f, 0.f, 0.f));
    }
    //         the quantised records, each from its own node's box and its children's boxes
    for (size_t k = 0; k < wide.size(); k++) {
        const WideNode& w = wide[k];
        const AABB& box = T.box[w.bin];
        const Vec3 ext = box.extent();
        const int ex = gridExponent(ext.x, box.lo.x, box.hi.x), ey = gridExponent(ext.y, box.lo.y, box.hi.y),
                  ez = gridExponent(ext.z, box.lo.z, box.hi.z);
        const float sx = ldexpf(1.f, ex), sy = ldexpf(1.f, ey), sz = ldexpf(1.f, ez);
        uint32_t planes[6][2] = { { 0, 0 }, { 0, 0 }, { 0, 0 }, { 0, 0 }, { 0, 0 }, { 0, 0 } };   // lo x y z, hi x y z
        uint32_t meta[2] = { 0, 0 }, imask = 0;
        int offset = 0, leafNo = 0;
        for (int s = 0; s < 8; s++) {
            const int c = w.child[s];
            if (c < 0) continue;
            const AABB& cb = T.box[c];
            const uint32_t q[6] = { gridStep(cb.lo.x, box.lo.x, sx, false), gridStep(cb.lo.y, box.lo.y, sy, false),
                                    gridStep(cb.lo.z, box.lo.z, sz, false), gridStep(cb.hi.x, box.lo.x, sx, true),
                                    gridStep(cb.hi.y, box.lo.y, sy, true), gridStep(cb.hi.z, box.lo.z, sz, true) };
            for (int p = 0; p < 6; p++) planes[p][s >> 2] |= q[p] << (8 * (s & 3));
            uint32_t m;
            if (w.isLeaf[s]) {
                const int cnt = leafTris[k][leafNo++];
                m = 0x80u | ((uint32_t)(cnt - 1) << 5) | (uint32_t)offset;
                offset += cnt;
            } else { m = 0x40u; imask |= 1u << s; }
            meta[s >> 2] |= m << (8 * (s & 3));
        }
        Vec4* o = &W.nodes[(size_t)nodeId[k] * 5];
        const uint32_t ew = (uint32_t)(ex + 127) | ((uint32_t)(ey + 127) << 8) | ((uint32_t)(ez + 127) << 16) | (imask << 24);
        o[0] = v4(box.lo.x, box.lo.y, box.lo.z, asFloatU(ew));
        o[1] = v4(asFloat(childBase[k]), asFloat(triBase[k]), asFloatU(meta[0]), asFloatU(meta[1]));
        o[2] = v4(asFloatU(planes[0][0]), asFloatU(planes[0][1]), asFloatU(planes[1][0]), asFloatU(planes[1][1]));
        o[3] = v4(asFloatU(planes[2][0]), asFloatU(planes[2][1]), asFloatU(planes[3][0]), asFloatU(planes[3][1]));
        o[4] = v4(asFloatU(planes[4][0]), asFloatU(planes[4][1]), asFloatU(planes[5][0]), asFloatU(planes[5][1]));
    }
}

// ------------------------------------------------------------------------------------------------ triangle packing
namespace {
// inverse of a 3x3 matrix by its cofactors (determinant by the first row); a singular matrix gives zeros
void inverse3(const double m[3][3], double inv[3][3]) {
    double cof[3][3];
    for (int r = 0; r < 3; r++)
        for (int c = 0; c < 3; c++) {
            const int r0 = r == 0 ? 1 : 0, r1 = r == 2 ? 1 : 2, c0 = c == 0 ? 1 : 0, c1 = c == 2 ? 1 : 2;
            // the signed minor, with the odd-position sign folded into the order of the difference (so an exact zero
            // stays +0 rather than becoming -0)
            const double p = m[r0][c0] * m[r1][c1], q = m[r0][c1] * m[r1][c0];
            cof[r][c] = (r + c) & 1 ? q - p : p - q;
        }
    const double det = m[0][0] * cof[0][0] + m[0][1] * cof[0][1] + m[0][2] * cof[0][2];
    for (int r = 0; r < 3; r++) for (int c = 0; c < 3; c++) inv[r][c] = 0.0;
    if (fabs(det) < 1e-300) return;
    const double s = 1.0 / det;
    for (int r = 0; r < 3; r++) for (int c = 0; c < 3; c++) inv[r][c] = cof[c][r] * s;
}
}  // namespace

void packWoop(const Mesh& m, const std::vector<int>& order, std::vector<Vec4>& rows) {
    const size_t n = order.size();
    rows.resize(3 * n);
    for (size_t k = 0; k < n; k++) {
        const Tri& t = m.f[order[k]];
        const Vec3 v0 = m.v[t.a], e1 = m.v[t.b] - v0, e2 = m.v[t.c] - v0, nn = cross(e1, e2);
        // the affine map taking the triangle to the unit triangle: rows of inverse([e1 e2 n]), translation -inv v0
        const double M[3][3] = { { e1.x, e2.x, nn.x }, { e1.y, e2.y, nn.y }, { e1.z, e2.z, nn.z } };
        double inv[3][3];
        inverse3(M, inv);
        for (int r = 0; r < 3; r++) {
            const double tr = -(inv[r][0] * v0.x + inv[r][1] * v0.y + inv[r][2] * v0.z);
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
