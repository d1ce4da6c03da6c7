"""Pins of the oracle's builders by independent restatements and brute force (CPU only).

* H-PLOC (src/accel/hploc.cu:56-136): an independent Python build from the sorted keys. The LBVH is derived top-down
  by Karras's radix-tree rule (a range splits where its highest differing key bit flips: the largest key XOR of
  neighbours), which yields the same tree as the reference's bottom-up Apetrei hand-over. Every LBVH node whose range
  exceeds 16 primitives, and the root, is reduced by PLOC iterations over the clusters of its two children (left
  survivors, then right survivors): nearest neighbour within 8 positions by the half area of the merged box, pairs
  ordered by (area bits, lower position, upper position), mutual pairs merge with the lower position leading, and the
  survivors keep their order, down to 16 clusters (1 at the root). The tree must equal the oracle's in canonical form.
* Binned SAH (src/accel/bvh_builder.cpp:14-68): at every internal node, every one of the 3 x (bins - 1) candidate
  splits is evaluated by brute force; the node's children must be the two sides of the first minimum (axis-major,
  then bin), or the median halves when all centroids coincide.
* CWBVH collapse (src/accel/bvh_tools.cpp:75-168): on trees of at most 12 leaves, every 8-wide collapse is enumerated;
  the oracle's wide tree must reach the minimum cost (c_node 1, c_prim 0.3, leaves of at most 3 triangles,
  src/accel/bvh_tools.h:19,27). Each node's per-axis exponent must be the smallest e with
  lo + 255 * 2^e >= hi, and each quantised child plane the tightest integer plane at that exponent.
"""
import sys

import numpy as np
import pytest

import pyoracle as po

T, R = 16, 8


# ------------------------------------------------------------------------------------------------------ H-PLOC
def _half_area_rows(lo, hi):
    """float32 half area of each row box, (dx dy + dy dz) + dz dx without fused multiply-adds."""
    d = (hi - lo).astype(np.float32)
    return ((d[:, 0] * d[:, 1]) + (d[:, 1] * d[:, 2])) + (d[:, 2] * d[:, 0])


def _lbvh_postorder(keys):
    """Internal LBVH nodes (L, R, split) in post-order, from the top-down radix-tree rule on the sorted keys."""
    n = len(keys)
    k = keys.astype(np.uint64)
    delta = k[:-1] ^ k[1:]
    out, st = [], [(0, n - 1, False)]
    while st:
        L, Rr, done = st.pop()
        if L == Rr:
            continue
        s = L + int(np.argmax(delta[L:Rr]))
        if done:
            out.append((L, Rr, s))
            continue
        st.append((L, Rr, True))
        st.append((s + 1, Rr, False))
        st.append((L, s, False))
    # the second visit of each node follows both children: post-order; recompute split for the emitted tuples
    return out


def _ploc_reduce(ids, boxes, target, new_node):
    """PLOC iterations on a window of cluster ids (boxes: id -> (lo, hi) float32[3]) down to `target` clusters."""
    ids = list(ids)
    while len(ids) > target:
        m = len(ids)
        lo = np.array([boxes[i][0] for i in ids], np.float32)
        hi = np.array([boxes[i][1] for i in ids], np.float32)
        best = [None] * m
        for r in range(1, R + 1):
            if r >= m:
                break
            a = np.arange(m - r)
            area = _half_area_rows(np.minimum(lo[a], lo[a + r]), np.maximum(hi[a], hi[a + r]))
            bits = area.view(np.uint32).astype(np.uint64)
            for x in range(m - r):
                key = (int(bits[x]) << 32) | (x << 5) | (x + r)
                for y in (x, x + r):
                    if best[y] is None or key < best[y]:
                        best[y] = key
        nn = []
        for y in range(m):
            pa, pb = (best[y] >> 5) & 31, best[y] & 31
            nn.append(pb if pa == y else pa)
        out = []
        merged = False
        for y in range(m):
            p = nn[y]
            mutual = nn[p] == y and p != y
            if mutual and y > p:
                continue
            if mutual:
                merged = True
                out.append(new_node(ids[y], ids[p]))
            else:
                out.append(ids[y])
        assert merged
        ids = out
    return ids


def hploc_replay(lo, hi):
    """Independent H-PLOC build: (children of each internal node, leaf primitive, box) from the oracle's pinned keys."""
    lo3, hi3 = lo[:, :3].astype(np.float32), hi[:, :3].astype(np.float32)
    n = len(lo3)
    _, keys, mb, ib = po.hp_keys(lo, hi)
    keys = np.sort(keys)
    prim = (keys & np.uint64((1 << ib) - 1)).astype(np.int64)
    boxes = {i: (lo3[prim[i]].copy(), hi3[prim[i]].copy()) for i in range(n)}
    kids = {}

    def new_node(a, b):
        x = n + len(kids)
        kids[x] = (a, b)
        boxes[x] = (np.minimum(boxes[a][0], boxes[b][0]), np.maximum(boxes[a][1], boxes[b][1]))
        return x

    if n == 1:
        return kids, prim, boxes, 0
    clusters = {}   # LBVH range start -> survivors of the merged node that starts there

    def survivors(L, Rr):
        if (L, Rr) in clusters:
            return clusters[(L, Rr)]
        return list(range(L, Rr + 1))   # not merged: its leaves in order (a range of at most 16)

    root = None
    for L, Rr, s in _lbvh_postorder(keys):
        is_root = L == 0 and Rr == n - 1
        if Rr - L + 1 > T or is_root:
            win = survivors(L, s) + survivors(s + 1, Rr)
            assert len(win) <= 2 * T
            clusters[(L, Rr)] = _ploc_reduce(win, boxes, 1 if is_root else T, new_node)
            if is_root:
                root = clusters[(L, Rr)][0]
    return kids, prim, boxes, root


def canonical(kids, prim, boxes, root, n):
    out, st = [], [(root, -1, 0)]
    while st:
        x, parent, side = st.pop()
        me = len(out)
        if parent >= 0:
            out[parent][2 + side] = me
        b = tuple(np.concatenate([boxes[x][0], boxes[x][1]]).astype(np.float32).view(np.uint32).tolist())
        if x < n:
            out.append([b, "leaf", int(prim[x]), -1])
        else:
            out.append([b, "node", -1, -1])
            st.append((kids[x][1], me, 1))
            st.append((kids[x][0], me, 0))
    return [tuple(r) for r in out]


def _boxes(n, seed, kind):
    rng = np.random.default_rng(seed)
    if kind == "uniform":
        c = rng.random((n, 3)).astype(np.float32)
        lo, hi = c - np.float32(0.004), c + rng.random((n, 3)).astype(np.float32) * np.float32(0.02)
    elif kind == "clustered":   # tight clusters: many equal Morton prefixes, deep LBVH chains
        centres = rng.random((max(1, n // 50), 3)).astype(np.float32)
        c = centres[rng.integers(len(centres), size=n)] + rng.normal(0, 0.002, (n, 3)).astype(np.float32)
        lo, hi = c - np.float32(0.001), c + np.float32(0.001)
    else:   # a heightfield's triangle boxes
        g = int(np.sqrt(n / 2)) + 1
        xs = np.linspace(0, 1, g, dtype=np.float32)
        X, Y = np.meshgrid(xs, xs)
        Z = (0.05 * rng.standard_normal((g, g))).astype(np.float32)
        v = np.stack([X.ravel(), Y.ravel(), Z.ravel()], 1)
        i = np.arange(g * g).reshape(g, g)
        a, b, c_, d = i[:-1, :-1].ravel(), i[:-1, 1:].ravel(), i[1:, :-1].ravel(), i[1:, 1:].ravel()
        t = np.concatenate([np.stack([a, b, d], 1), np.stack([a, d, c_], 1)])[:n]
        lo, hi = v[t].min(1), v[t].max(1)
    z = np.zeros((len(lo), 1), np.float32)
    return np.concatenate([lo, z], 1).astype(np.float32), np.concatenate([hi, z], 1).astype(np.float32)


@pytest.mark.parametrize("n,kind", [(2, "uniform"), (33, "uniform"), (300, "uniform"), (2000, "clustered"),
                                    (5000, "heightfield"), (20000, "uniform")])
def test_hploc_oracle_equals_independent_replay_of_every_window(n, kind):
    lo, hi = _boxes(n, 7 + n, kind)
    sys.setrecursionlimit(10000)
    kids, prim, boxes, root = hploc_replay(lo, hi)
    assert len(kids) == n - 1
    nodes, oroot, _ = po.hp_build(lo, hi)
    assert canonical(kids, prim, boxes, root, n) == po.hp_canonical(nodes, n, oroot)


def test_hploc_replay_detects_a_changed_radius():
    """The replay is sensitive to the merge rule: with a search radius of 4 instead of 8 the trees differ."""
    global R
    lo, hi = _boxes(600, 3, "uniform")
    nodes, oroot, _ = po.hp_build(lo, hi)
    saved = R
    try:
        R = 4
        kids, prim, boxes, root = hploc_replay(lo, hi)
    finally:
        R = saved
    assert canonical(kids, prim, boxes, root, 600) != po.hp_canonical(nodes, 600, oroot)


# ------------------------------------------------------------------------------------------------------ binned SAH
def _area(lo, hi):
    """AABB::area in float32: 0 for an empty box, else 2 ((dx dy + dy dz) + dz dx)."""
    if hi[0] < lo[0]:
        return np.float32(0)
    d = (hi - lo).astype(np.float32)
    return np.float32(2) * ((d[0] * d[1] + d[1] * d[2]) + d[2] * d[0])


def _subtree_prims(tree, prim, k):
    out, st = [], [k]
    while st:
        x = st.pop()
        if tree[x, 3] > 0:
            out.extend(prim[tree[x, 2]:tree[x, 2] + tree[x, 3]].tolist())
        else:
            st.extend([tree[x, 1], tree[x, 0]])
    return out


def _mesh(n_boxes, seed):
    rng = np.random.default_rng(seed)
    cube = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0], [0, 0, 1], [1, 0, 1], [1, 1, 1], [0, 1, 1]], np.float32)
    faces = np.array([[0, 1, 2], [0, 2, 3], [4, 6, 5], [4, 7, 6], [0, 4, 5], [0, 5, 1], [1, 5, 6], [1, 6, 2],
                      [2, 6, 7], [2, 7, 3], [3, 7, 4], [3, 4, 0]], np.int32)
    v, t = [], []
    for b in range(n_boxes):
        c = rng.random(3).astype(np.float32)
        e = rng.uniform(0.01, 0.08, 3).astype(np.float32)
        v.append((c + (cube - 0.5) * e).astype(np.float32))
        t.append(faces + 8 * b)
    return np.concatenate(v), np.concatenate(t)


@pytest.mark.parametrize("bins,leaf_max,n_boxes", [(16, 4, 60), (32, 4, 60), (16, 1, 40), (32, 2, 25)])
def test_sah_splits_are_the_first_minimum_of_all_candidates(bins, leaf_max, n_boxes):
    v, t = _mesh(n_boxes, bins + leaf_max)
    tb_lo, tb_hi = v[t].min(1), v[t].max(1)
    cen = (tb_lo + tb_hi) * np.float32(0.5)
    out = po.tree_arrays(po.build_sah(v, t, bins, leaf_max))
    tree, prim = out["tree"], out["prim"]
    checked = 0
    for k in range(len(tree)):
        P = np.array(sorted(_subtree_prims(tree, prim, k)))
        if tree[k, 3] > 0:
            assert len(P) <= leaf_max
            continue
        assert len(P) > leaf_max
        left = set(_subtree_prims(tree, prim, tree[k, 0]))
        clo, chi = cen[P].min(0), cen[P].max(0)
        ext = (chi - clo).astype(np.float32)
        best = (np.float32(np.finfo(np.float32).max), -1, -1)
        for axis in range(3):
            e = ext[axis]
            if e <= np.float32(1e-12):
                continue
            scale = np.float32(bins) / e
            b = ((cen[P, axis] - clo[axis]) * scale).astype(np.float32).astype(np.int64).clip(0, bins - 1)
            for split in range(bins - 1):   # brute force: each candidate's two sides from scratch
                L, Rr = P[b <= split], P[b > split]
                if len(L) == 0 or len(Rr) == 0:
                    continue
                cost = (_area(tb_lo[L].min(0), tb_hi[L].max(0)) * np.float32(len(L)) +
                        _area(tb_lo[Rr].min(0), tb_hi[Rr].max(0)) * np.float32(len(Rr)))
                if cost < best[0]:
                    best = (cost, axis, split)
        if best[1] < 0:   # all centroids coincide: a median split of the node's own order
            assert len(left) == len(P) // 2
        else:
            e = ext[best[1]]
            b = ((cen[P, best[1]] - clo[best[1]]) * (np.float32(bins) / e)).astype(np.float32).astype(np.int64).clip(0, bins - 1)
            want = set(P[b <= best[2]].tolist())
            if len(want) not in (0, len(P)):
                assert left == want, (k, best)
        checked += 1
    assert checked > 10


# ------------------------------------------------------------------------------------------------------ CWBVH
def _area64(lo, hi):
    d = (hi - lo).astype(np.float64)
    return 2.0 * (d[0] * d[1] + d[1] * d[2] + d[2] * d[0])


def _cwbvh_min_cost(tree, box, p_max=3, c_node=1.0, c_prim=0.3):
    """Minimum cost of any 8-wide collapse of a binary tree, by enumerating every cut: a wide node at binary node x
    has as children any set of at most 8 binary descendants that covers x's subtree (each a leaf of at most p_max
    primitives or a wide node); cost = sum over wide nodes of A c_node + sum over leaves of A count c_prim."""
    memo_f, memo_cuts = {}, {}
    count = {}

    def cnt(x):
        if x not in count:
            count[x] = int(tree[x, 3]) if tree[x, 3] > 0 else cnt(tree[x, 0]) + cnt(tree[x, 1])
        return count[x]

    def cuts(x):   # every cover of x's subtree by descendants (x itself included), as tuples of size <= 8
        if x in memo_cuts:
            return memo_cuts[x]
        res = {(x,)}
        if tree[x, 3] == 0:
            for a in cuts(tree[x, 0]):
                for b in cuts(tree[x, 1]):
                    if len(a) + len(b) <= 8:
                        res.add(a + b)
        memo_cuts[x] = res
        return res

    def f(x):   # best cost of subtree x as ONE child (a leaf or a wide node)
        if x in memo_f:
            return memo_f[x]
        A = _area64(box[x, :3], box[x, 3:])
        best = A * cnt(x) * c_prim if cnt(x) <= p_max else np.inf
        if tree[x, 3] == 0:
            opts = []
            for a in cuts(tree[x, 0]):
                for b in cuts(tree[x, 1]):
                    if len(a) + len(b) <= 8:
                        opts.append(sum(f(c) for c in a + b))
            best = min(best, A * c_node + min(opts))
        memo_f[x] = best
        return best

    # the root is always a wide node unless the whole tree is one leaf
    return f(0), cnt


def _wide_cost(nodes, order_len, v, t, order, c_node=1.0, c_prim=0.3):
    """Cost of the oracle's wide tree (A c_node per wide node, A count c_prim per leaf slot) from its own encoding: wide-node
    boxes are their stored origin + the exact union of the triangles below them."""
    tb_lo, tb_hi = v[t].min(1), v[t].max(1)
    ints = nodes.view(np.int32).reshape(-1, 20)
    u = nodes.view(np.uint32).reshape(-1, 20)

    def tris_of(k):
        out = []
        ew = u[k, 3]
        imask = ew >> 24
        child_base, tri_base = ints[k, 4], ints[k, 5]
        meta = [(u[k, 6 + s // 4] >> (8 * (s % 4))) & 255 for s in range(8)]
        for s in range(8):
            m = meta[s]
            if m & 0x80:
                out.extend(range(tri_base + (m & 31), tri_base + (m & 31) + ((m >> 5) & 3) + 1))
            elif imask & (1 << s):
                out.extend(tris_of(child_base + bin(imask & ((1 << s) - 1)).count("1")))
        return out

    total = 0.0
    for k in range(len(ints)):
        tr = order[np.array(tris_of(k))]
        lo, hi = tb_lo[tr].min(0), tb_hi[tr].max(0)
        total += _area64(lo, hi) * c_node
        meta = [(u[k, 6 + s // 4] >> (8 * (s % 4))) & 255 for s in range(8)]
        tri_base = ints[k, 5]
        for s in range(8):
            m = meta[s]
            if m & 0x80:
                leaf = order[tri_base + (m & 31): tri_base + (m & 31) + ((m >> 5) & 3) + 1]
                total += _area64(tb_lo[leaf].min(0), tb_hi[leaf].max(0)) * len(leaf) * c_prim
    return total


@pytest.mark.parametrize("seed", [1, 2, 3])
@pytest.mark.parametrize("n_tris", [5, 8, 9, 12])
def test_cwbvh_collapse_reaches_the_minimum_over_all_collapses(seed, n_tris):
    rng = np.random.default_rng(seed * 31 + n_tris)
    # n_tris random small triangles spread over the unit cube (a tree of n_tris binary leaves with leaf_max 1)
    c = rng.random((n_tris, 3)).astype(np.float32)
    v = np.concatenate([c, c + np.float32(0.03) * rng.random((n_tris, 3)).astype(np.float32),
                        c + np.float32(0.03) * rng.random((n_tris, 3)).astype(np.float32)]).astype(np.float32)
    t = np.stack([np.arange(n_tris), np.arange(n_tris) + n_tris, np.arange(n_tris) + 2 * n_tris], 1).astype(np.int32)
    tr = po.build_sah(v, t, 16, 1)
    ta = po.tree_arrays(tr)
    out = po._layout(tr, 8)
    po.lib().orc_tree_free(po.ctypes.byref(tr))
    best, _ = _cwbvh_min_cost(ta["tree"], ta["tree_box"])
    got = _wide_cost(out["nodes"], len(out["order"]), v, t, out["order"])
    assert got == pytest.approx(best, rel=1e-5)


def test_cwbvh_exponents_are_smallest_and_planes_tightest():
    v, t = _mesh(80, 5)
    b = po.build(v, t, variant=8, bins=16)
    nodes, order = b["nodes"], b["order"]
    tb_lo, tb_hi = v[t].min(1), v[t].max(1)
    u = nodes.view(np.uint32).reshape(-1, 20)
    ints = nodes.view(np.int32).reshape(-1, 20)

    def plane_bytes(k, w):
        return [(u[k, 8 + w] >> (8 * j)) & 255 for j in range(4)] + [(u[k, 9 + w] >> (8 * j)) & 255 for j in range(4)]

    def child_tris(k, s):
        ew = u[k, 3]
        imask = ew >> 24
        meta = (u[k, 6 + s // 4] >> (8 * (s % 4))) & 255
        if meta & 0x80:
            tb = ints[k, 5] + (meta & 31)
            return list(range(tb, tb + ((meta >> 5) & 3) + 1))
        c = ints[k, 4] + bin(imask & ((1 << s) - 1)).count("1")
        return [x for s2 in range(8) if ((u[c, 6 + s2 // 4] >> (8 * (s2 % 4))) & 255) for x in child_tris(c, s2)]

    planes_checked = 0
    for k in range(len(u)):
        tris = [x for s in range(8) if ((u[k, 6 + s // 4] >> (8 * (s % 4))) & 255) for x in child_tris(k, s)]
        lo = tb_lo[order[tris]].min(0)
        hi = tb_hi[order[tris]].max(0)
        np.testing.assert_array_equal(nodes.reshape(-1, 20)[k, :3], lo)
        for a in range(3):
            e = int((u[k, 3] >> (8 * a)) & 255) - 127
            if not hi[a] > lo[a]:
                continue
            # smallest exponent reaching the box's upper side with 255 steps (float32, as the encoder evaluates it)
            assert np.float32(lo[a]) + np.float32(255) * np.float32(2.0 ** e) >= hi[a]
            assert np.float32(lo[a]) + np.float32(255) * np.float32(2.0 ** (e - 1)) < hi[a]
            step = 2.0 ** e
            qlo, qhi = plane_bytes(k, 2 * a), plane_bytes(k, 6 + 2 * a)
            for s in range(8):
                if not ((u[k, 6 + s // 4] >> (8 * (s % 4))) & 255):
                    continue
                ct = order[child_tris(k, s)]
                clo, chi = float(tb_lo[ct, a].min()), float(tb_hi[ct, a].max())
                # conservative, and the tightest integer plane: one step further in would cut the child's box
                assert lo[a] + qlo[s] * step <= clo and (qlo[s] == 255 or lo[a] + (qlo[s] + 1) * step > clo)
                assert lo[a] + qhi[s] * step >= chi and (qhi[s] == 0 or lo[a] + (qhi[s] - 1) * step < chi)
                planes_checked += 2
    assert planes_checked > 500
