"""GPU compressed-BVH8 collapse of H-PLOC trees (device/collapse.hip) against the host's (treeFromHploc +
collapseCWBVH, itself bit-exact with the oracle): the 80-byte nodes and the leaf primitive order must be identical,
bit for bit."""
import numpy as np
import pytest

import cwp
import pyoracle as po

pytestmark = pytest.mark.gpu


def boxes(n, seed, spread=0.02):
    rng = np.random.default_rng(seed)
    c = rng.random((n, 3)).astype(np.float32)
    return c - np.float32(0.005), c + rng.random((n, 3)).astype(np.float32) * np.float32(spread)


def tris_of(lo, hi):
    """a triangle per box (the layout depends only on the H-PLOC boxes; the triangles satisfy the host API)"""
    n = len(lo)
    v = np.concatenate([lo, hi, np.stack([lo[:, 0], hi[:, 1], hi[:, 2]], 1)]).astype(np.float32)
    t = np.stack([np.arange(n), np.arange(n) + n, np.arange(n) + 2 * n], 1).astype(np.int32)
    return v, t


def compare(lo, hi, g=None):
    n = len(lo)
    if g is None:
        g, root, _ = cwp.hploc_build(lo, hi)
    root = 2 * n - 2 if n > 1 else 0
    v, t = tris_of(lo, hi)
    host = cwp.bvh_from_hploc(v, t, g, root, 8)
    nodes, order, depth, ms = cwp.cwbvh_from_hploc_gpu(g, n)
    hn = host["nodes"].reshape(-1, 20)
    assert nodes.shape == hn.shape, (nodes.shape, hn.shape)
    np.testing.assert_array_equal(nodes.view(np.uint32), hn.view(np.uint32))
    np.testing.assert_array_equal(order, host["order"])
    assert sorted(order.tolist()) == list(range(n))
    assert depth >= 1 and ms > 0


@pytest.mark.parametrize("n", [1, 2, 3, 4, 9, 100, 5000, 70000])
def test_collapse_matches_host_random_boxes(n):
    print("case collapse n=%d" % n, flush=True)
    lo, hi = boxes(n, 100 + n)
    compare(lo, hi)


def test_collapse_matches_host_mesh_and_degenerate_boxes():
    print("case collapse terrain", flush=True)
    s = cwp.Setup(["--gen", "terrain", "--gen_grid", "121", "--gen_boxes", "50", "--width", "8", "--height", "8"])
    lo, hi = po.tri_boxes(s.array(cwp.ARR_VERTICES), s.array(cwp.ARR_TRIANGLES))
    compare(lo[:, :3].astype(np.float32), hi[:, :3].astype(np.float32))
    print("case collapse equal boxes", flush=True)
    same = np.tile(np.array([[0.5, 0.5, 0.5]], np.float32), (300, 1))
    compare(same, same + np.float32(0.1))
    print("case collapse flat boxes", flush=True)
    lo, hi = boxes(2000, 7)
    lo[:, 2] = 0.25; hi[:, 2] = 0.25   # zero extent in z
    compare(lo, hi)


def test_collapse_rejects_a_malformed_tree():
    lo, hi = boxes(64, 3)
    g, root, _ = cwp.hploc_build(lo, hi)
    bad = g.copy()
    bad.view(np.int32)[70, 3] = 70   # an internal node naming itself as a child
    with pytest.raises(cwp.CwpError, match="not a tree"):
        cwp.cwbvh_from_hploc_gpu(bad, 64)


def caterpillar(n):
    """H-PLOC-layout nodes of a chain: internal node n + k joins the previous internal node (leaf 0 for k = 0) and leaf
    k + 1; leaves are boxes along x, so the tree is as deep as it is large"""
    lo = np.stack([np.arange(n), np.zeros(n), np.zeros(n)], 1).astype(np.float32)
    hi = lo + np.float32(0.5)
    g = np.zeros((2 * n - 1, 8), np.float32)
    gi = g.view(np.int32)
    g[:n, 0:3] = lo; g[:n, 4:7] = hi
    gi[:n, 3] = np.arange(n); gi[:n, 7] = -1
    for k in range(n - 1):
        a, b = (0 if k == 0 else n + k - 1), k + 1
        g[n + k, 0:3] = np.minimum(g[a, 0:3], g[b, 0:3]); g[n + k, 4:7] = np.maximum(g[a, 4:7], g[b, 4:7])
        gi[n + k, 3] = a; gi[n + k, 7] = b
    return lo, hi, g


@pytest.mark.parametrize("n", [300, 1000])
def test_collapse_matches_host_deep_chain(n):
    """a binary chain collapses into a wide tree tens of levels deep: the device level loop runs several batches of
    levels between its host checks and must still give the host's layout"""
    print("case collapse chain n=%d" % n, flush=True)
    lo, hi, g = caterpillar(n)
    compare(lo, hi, g)
    _, _, depth, _ = cwp.cwbvh_from_hploc_gpu(g, n)
    assert depth > 8
