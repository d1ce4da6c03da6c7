"""GPU parity tests of the H-PLOC builder (device/hploc.hip) against the oracle's restatement (oracle/hploc.c).

Internal node ids come from an atomic counter, so their numbering depends on scheduling; the tree does not. Trees are
compared in a canonical pre-order form (boxes bit for bit, leaves by primitive id, children in left/right order).
"""
import numpy as np
import pytest

import cwp
import pyoracle as po

pytestmark = pytest.mark.gpu


def boxes(n, seed, spread=0.02):
    rng = np.random.default_rng(seed)
    c = rng.random((n, 3)).astype(np.float32)
    return c - np.float32(0.005), c + rng.random((n, 3)).astype(np.float32) * np.float32(spread)


def check_tree(nodes, root, lo, hi):
    n = len(lo)
    ints = nodes.view(np.int32)
    assert root == (2 * n - 2 if n > 1 else 0)
    assert sorted(ints[:n, 3].tolist()) == list(range(n))
    if n > 1:
        kids = ints[n:, [3, 7]]
        assert (kids < np.arange(n, 2 * n - 1)[:, None]).all()
        assert sorted(kids.ravel().tolist()) == list(range(2 * n - 2))


# 16384 is the largest count sorted inside one workgroup; larger counts use the multi-workgroup radix passes
@pytest.mark.parametrize("n", [1, 2, 3, 17, 100, 1000, 16384, 16385, 70000])
def test_hploc_matches_oracle(n):
    print("case hploc n=%d" % n, flush=True)
    lo, hi = boxes(n, n)
    g, groot, ms = cwp.hploc_build(lo, hi)
    o, oroot, _ = po.hp_build(lo, hi)
    check_tree(g, groot, lo, hi)
    assert po.hp_canonical(g, n, groot) == po.hp_canonical(o, n, oroot)
    # leaves are the primitives in Morton order: identical, not only up to numbering
    np.testing.assert_array_equal(g[:n].view(np.uint32), o[:n].view(np.uint32))
    assert ms > 0


@pytest.mark.parametrize("n", [1, 2, 100, 4097, 16385, 70000])
def test_hploc_30bit_matches_oracle(n):
    print("case hploc30 n=%d" % n, flush=True)
    lo, hi = boxes(n, n + 7)
    g, groot, _ = cwp.hploc_build(lo, hi, cwp.MORTON_30)
    o, oroot, _ = po.hp_build(lo, hi, 30)
    check_tree(g, groot, lo, hi)
    assert po.hp_canonical(g, n, groot) == po.hp_canonical(o, n, oroot)
    np.testing.assert_array_equal(g[:n].view(np.uint32), o[:n].view(np.uint32))


@pytest.mark.parametrize("morton", [0, 30])
def test_hploc_200k_triangle_mesh_matches_oracle(morton):
    """The largest BLAS of the two-level configuration (about 200k triangles), both Morton modes."""
    print("case hploc mesh 200k morton=%d" % morton, flush=True)
    s = cwp.Setup(["--gen", "terrain", "--gen_grid", "317", "--gen_boxes", "0", "--width", "8", "--height", "8"])
    lo, hi = po.tri_boxes(s.array(cwp.ARR_VERTICES), s.array(cwp.ARR_TRIANGLES))
    n = len(lo)
    assert n == 2 * 316 * 316
    g, groot, _ = cwp.hploc_build(lo, hi, morton)
    o, oroot, _ = po.hp_build(lo, hi, morton)
    check_tree(g, groot, lo, hi)
    assert po.hp_canonical(g, n, groot) == po.hp_canonical(o, n, oroot)
    np.testing.assert_array_equal(g[:n].view(np.uint32), o[:n].view(np.uint32))


def test_hploc_on_mesh_triangles_and_duplicate_centroids():
    s = cwp.Setup(["--gen", "terrain", "--gen_grid", "121", "--gen_boxes", "50", "--width", "8", "--height", "8"])
    lo, hi = po.tri_boxes(s.array(cwp.ARR_VERTICES), s.array(cwp.ARR_TRIANGLES))
    g, groot, _ = cwp.hploc_build(lo, hi)
    o, oroot, _ = po.hp_build(lo, hi)
    assert po.hp_canonical(g, len(lo), groot) == po.hp_canonical(o, len(lo), oroot)
    # all boxes equal: every Morton code is the same, the primitive index alone orders the keys
    same = np.tile(np.array([[0.5, 0.5, 0.5, 0.0]], np.float32), (300, 1))
    g, groot, _ = cwp.hploc_build(same, same + np.float32(0.1))
    o, oroot, _ = po.hp_build(same, same + np.float32(0.1))
    assert po.hp_canonical(g, 300, groot) == po.hp_canonical(o, 300, oroot)


def test_hploc_device_api_and_bvh2_layout():
    n = 5000
    lo, hi = boxes(n, 11)
    dlo, dhi = cwp.DeviceArray.from_host(po._box4(lo)), cwp.DeviceArray.from_host(po._box4(hi))
    h = cwp.ctypes.c_void_p()
    cwp._check(cwp.lib().cwp_hploc_create(1000, cwp.ctypes.byref(h)))   # grows on demand
    try:
        root = cwp.ctypes.c_int()
        for _ in range(2):   # the builder is reusable
            cwp._check(cwp.lib().cwp_hploc_build(h, dlo.ptr, dhi.ptr, n, None, cwp.ctypes.byref(root)))
        ptr = cwp.ctypes.c_void_p()
        cwp._check(cwp.lib().cwp_hploc_nodes(h, cwp.ctypes.byref(ptr)))
        nodes = np.zeros((2 * n - 1, 8), np.float32)
        cwp._check(cwp.lib().cwp_memcpy(nodes.ctypes.data, ptr.value, nodes.nbytes, cwp.CWP_COPY_D2H))
        check_tree(nodes, root.value, lo, hi)
        o, oroot, _ = po.hp_build(lo, hi)
        assert po.hp_canonical(nodes, n, root.value) == po.hp_canonical(o, n, oroot)
        out = cwp.DeviceArray((n - 1) * 64)
        cwp._check(cwp.lib().cwp_hploc_to_bvh2(ptr.value, n, out.ptr, None))
        bvh2 = out.to_host(np.float32, (n - 1, 16))
        np.testing.assert_array_equal(bvh2.view(np.uint32), po.hp_to_bvh2(nodes, n).view(np.uint32))
    finally:
        cwp.lib().cwp_hploc_destroy(h)


@pytest.mark.parametrize("variant,gpu_collapse", [(8, 1), (8, 0), (2, 1)])
def test_builder_hploc_scene_matches_oracle_and_renders(variant, gpu_collapse):
    """--builder hploc: GPU H-PLOC over the triangle boxes, then the layout: for BVH8 the CWBVH collapse on the GPU
    (default) or on the host, for BVH2 the host SAH leaf collapse. The layout equals the oracle's build from the
    oracle's own H-PLOC restatement, bit for bit, and the image matches the oracle's."""
    print("case builder hploc bvh%d gpu_collapse %d" % (variant, gpu_collapse), flush=True)
    args = ["--gen", "boxes", "--width", "64", "--height", "64", "--builder", "hploc", "--variant", "bvh%d" % variant,
            "--batch", "8", "--spp", "8", "--bounces", "3", "--gpu_collapse", str(gpu_collapse)]
    s = cwp.Setup(args)
    v, t = s.array(cwp.ARR_VERTICES), s.array(cwp.ARR_TRIANGLES)
    o = po.build(v, t, variant, leaf_max=4, builder="hploc")
    np.testing.assert_array_equal(s.array(cwp.ARR_NODES).view(np.uint32), o["nodes"].view(np.uint32))
    np.testing.assert_array_equal(s.array(cwp.ARR_ORDER), o["order"])
    r = cwp.Renderer(s)
    st = r.render(8)
    img = r.accum()
    j = r.frame_json(st)
    assert j["builder"] == "hploc" and j["bvh_tree_ms"] > 0
    sref = cwp.Setup([a for a in args if a not in ("--builder", "hploc")])   # (--gpu_collapse is ignored with sah)
    ref = cwp.Renderer(sref)
    ref.render(8)
    # a different tree, the same nearest hits: the images agree (ties between coincident faces may pick another face)
    a, b = img[..., :3] / 8, ref.accum()[..., :3] / 8
    assert np.mean(np.all(a == b, axis=-1)) > 0.99
