"""Pins of the oracle (the CPU restatement used as the parity checker) that do not rely on the product.

Each test checks an oracle function against something its formula cannot agree with by construction: a closed form,
an invariant, brute force, or a statistical law. What each function is pinned by:

  hash32                 the published inverse of Wellons' lowbias32 (hash-prospector), whose constants are the
                         modular inverses of the forward ones: inverse(hash32(x)) = x on 69,635 inputs
  rng4                   bit balance, uniformity and independence statistics (its composition of hashes is pinned by
                         statistics only: the reference ships no RNG vectors and cannot be built here)
  Blackman-Harris table  closed-form CDF of the window (integral of the cosine series)
  warps                  unit length, hemisphere / cone support, exact moments (E[cos] = 2/3; cone cos uniform)
  SAH builder            primitives exactly once, boxes = union of contents, leaf sizes, a separable two-cluster split
  CWBVH / BVH2 layouts   dequantised child boxes contain the exact boxes, leaves <= 3 triangles, traversal = brute force
  traversal              closest hit = brute force over all triangles (3 triangle tests x 2 layouts), any hit =
                         (closest hit exists); watertight test has no gaps on shared vertices of a grid
  legacy shading         closed forms: Lambert plane under a constant sky gives albedo * sky exactly; with a sun,
                         albedo * (sky + E cos / pi) within the sun-cone spread; sky pixels = sky
"""
import math

import numpy as np
import pytest

import pyoracle as po


# ------------------------------------------------------------------------------------------------------------ RNG
def test_rng_uniform_and_independent():
    n = 20000
    u = np.array([po.rng4(i, 7, 3, 2, 1000) for i in range(n)])
    assert u.min() >= 0.0 and u.max() < 1.0
    assert abs(u.mean() - 0.5) < 0.01
    assert abs(u.var() - 1.0 / 12.0) < 0.005
    # 24-bit resolution: every value is a multiple of 2^-24
    assert np.all((u * 16777216.0) == np.floor(u * 16777216.0))
    # neighbouring dimensions and samples are uncorrelated
    v = np.array([po.rng4(i, 7, 3, 3, 1000) for i in range(n)])
    w = np.array([po.rng4(i, 8, 3, 2, 1000) for i in range(n)])
    assert abs(np.corrcoef(u, v)[0, 1]) < 0.03
    assert abs(np.corrcoef(u, w)[0, 1]) < 0.03


def lowbias32_inverse(x):
    """The published inverse of lowbias32 (C. Wellons, hash-prospector, 2018): every step of the hash undone in
    reverse order, the multipliers replaced by their inverses mod 2^32."""
    x = np.asarray(x, dtype=np.uint64)
    m = np.uint64(0xFFFFFFFF)
    x ^= x >> np.uint64(16)
    x = (x * np.uint64(0x43021123)) & m
    x ^= (x >> np.uint64(15)) ^ (x >> np.uint64(30))
    x = (x * np.uint64(0x1D69E2A5)) & m
    x ^= x >> np.uint64(16)
    return x


def test_hash_is_lowbias32():
    """The hash is a bijection undone by lowbias32's published inverse: a wrong shift or multiplier would not be."""
    assert (0x846CA68B * 0x43021123) % (1 << 32) == 1 and (0x7FEB352D * 0x1D69E2A5) % (1 << 32) == 1
    rng = np.random.default_rng(3)
    xs = np.concatenate([np.arange(4096), rng.integers(0, 1 << 32, 1 << 16, dtype=np.uint64),
                         [0xFFFFFFFF, 0x80000000, 0x7FFFFFFF]]).astype(np.uint64)
    hs = np.array([po.hash32(int(x)) for x in xs], dtype=np.uint64)
    assert np.array_equal(lowbias32_inverse(hs), xs)
    assert po.hash32(0) == 0   # every step maps 0 to 0


def test_hash_bit_balance():
    hs = np.array([po.hash32(i) for i in range(4096)], dtype=np.uint64)
    bits = np.array([(hs >> b) & 1 for b in range(32)], dtype=np.float64)
    assert np.all(np.abs(bits.mean(axis=1) - 0.5) < 0.05)
    # avalanche: flipping one input bit flips about half the output bits
    flips = [bin(po.hash32(i) ^ po.hash32(i ^ 1)).count("1") for i in range(2000)]
    assert abs(np.mean(flips) - 16) < 1.0


# ------------------------------------------------------------------------------------------- Blackman-Harris table
def _bh_cdf(x):
    a0, a1, a2, a3 = 0.35875, 0.48829, 0.14128, 0.01168
    two = 2 * math.pi
    val = a0 * x - a1 * math.sin(two * x) / two + a2 * math.sin(2 * two * x) / (2 * two) - a3 * math.sin(3 * two * x) / (3 * two)
    return val / a0


def test_blackman_harris_table_closed_form():
    width = 1.5
    tab = po.blackman_harris_table(64, width)
    assert tab[0] == pytest.approx(-width / 2, abs=1e-6) and tab[-1] == pytest.approx(width / 2, abs=1e-6)
    assert np.all(np.diff(tab) > 0)
    np.testing.assert_allclose(tab + tab[::-1], 0.0, atol=2e-6)   # symmetric window -> antisymmetric inverse CDF
    for k in range(65):
        x = tab[k] / width + 0.5
        assert _bh_cdf(x) == pytest.approx(k / 64, abs=2e-4)


# ---------------------------------------------------------------------------------------------------------- warps
def test_cosine_hemisphere_moments():
    rng = np.random.default_rng(3)
    n = np.array([0.3, -0.5, 0.8], dtype=np.float32); n /= np.linalg.norm(n)
    d = np.array([po.cosine_hemisphere(n, a, b) for a, b in rng.random((20000, 2), dtype=np.float32)])
    np.testing.assert_allclose(np.linalg.norm(d, axis=1), 1.0, atol=1e-6)
    c = d @ n
    assert c.min() >= -1e-6
    assert c.mean() == pytest.approx(2.0 / 3.0, abs=0.01)        # E[cos] under pdf cos/pi
    assert (c ** 2).mean() == pytest.approx(0.5, abs=0.01)       # E[cos^2] = 1/2


def test_uniform_cone_distribution():
    rng = np.random.default_rng(4)
    axis = np.array([0.6, -0.4, 0.69], dtype=np.float32); axis /= np.linalg.norm(axis)
    cmax = 0.9
    d = np.array([po.uniform_cone(axis, cmax, a, b) for a, b in rng.random((20000, 2), dtype=np.float32)])
    c = d @ axis
    assert c.min() >= cmax - 1e-5 and c.max() <= 1 + 1e-6
    assert c.mean() == pytest.approx((1 + cmax) / 2, abs=0.002)  # cos uniform on [cmax, 1]
    assert c.var() == pytest.approx((1 - cmax) ** 2 / 12, rel=0.05)


# ---------------------------------------------------------------------------------------------------- BVH builder
def random_soup(n, seed, spread=1.0, size=0.05):
    rng = np.random.default_rng(seed)
    c = rng.random((n, 3)).astype(np.float32) * spread
    v = (c[:, None, :] + rng.normal(scale=size, size=(n, 3, 3)).astype(np.float32)).reshape(-1, 3)
    return v.astype(np.float32), np.arange(3 * n, dtype=np.int32).reshape(-1, 3)


def grid_mesh(g, seed=0, amp=0.05):
    rng = np.random.default_rng(seed)
    xs, ys = np.meshgrid(np.linspace(0, 1, g), np.linspace(0, 1, g))
    z = amp * rng.random((g, g))
    v = np.stack([xs, ys, z], -1).reshape(-1, 3).astype(np.float32)
    t = []
    for j in range(g - 1):
        for i in range(g - 1):
            a = j * g + i
            t += [(a, a + 1, a + g + 1), (a, a + g + 1, a + g)]
    return v, np.array(t, dtype=np.int32)


def tri_boxes(v, t):
    p = v[t]
    return p.min(axis=1), p.max(axis=1)


@pytest.mark.parametrize("bins,leaf", [(16, 4), (32, 4), (16, 1)])
def test_sah_tree_invariants(bins, leaf):
    v, t = random_soup(3000, 1)
    a = po.tree_arrays(po.build_sah(v, t, bins, leaf))
    node, box, prim = a["tree"], a["tree_box"], a["prim"]
    assert sorted(prim.tolist()) == list(range(len(t)))
    lo, hi = tri_boxes(v, t)
    seen = np.zeros(len(t), dtype=int)
    for i, (l, r, first, cnt) in enumerate(node):
        if cnt > 0:
            assert cnt <= leaf
            ids = prim[first:first + cnt]
            seen[ids] += 1
            np.testing.assert_array_equal(box[i, :3], lo[ids].min(0))
            np.testing.assert_array_equal(box[i, 3:], hi[ids].max(0))
        else:
            assert l > i and r > i                                   # pre-order ids
            np.testing.assert_array_equal(box[i, :3], np.minimum(box[l, :3], box[r, :3]))
            np.testing.assert_array_equal(box[i, 3:], np.maximum(box[l, 3:], box[r, 3:]))
    assert np.all(seen == 1)


def test_sah_separates_two_clusters():
    v1, t1 = random_soup(200, 2, spread=0.2)
    v2, t2 = random_soup(200, 3, spread=0.2)
    v2 = v2 + np.float32(5.0)
    v = np.concatenate([v1, v2]); t = np.concatenate([t1, t2 + len(v1)])
    a = po.tree_arrays(po.build_sah(v, t, 16, 4))
    l, r, _, _ = a["tree"][0]

    def leaves(n):
        nd = a["tree"][n]
        if nd[3] > 0:
            return set(a["prim"][nd[2]:nd[2] + nd[3]].tolist())
        return leaves(nd[0]) | leaves(nd[1])
    assert leaves(l) == set(range(200)) and leaves(r) == set(range(200, 400))


def _decode_cwbvh_child_boxes(nodes):
    """Yields (node, slot, decoded lo, decoded hi, meta) for every occupied slot of a CWBVH node array (n, 20)."""
    u = nodes.view(np.uint32)
    for i in range(len(nodes)):
        ew = int(u[i, 3])
        e = [(ew >> s) & 255 for s in (0, 8, 16)]
        scale = [2.0 ** (x - 127) for x in e]
        org = nodes[i, :3].astype(np.float64)
        meta = int(u[i, 6]) | (int(u[i, 7]) << 32)
        q = [int(u[i, 8 + 2 * k]) | (int(u[i, 9 + 2 * k]) << 32) for k in range(6)]
        for s in range(8):
            m = (meta >> (8 * s)) & 255
            if m == 0:
                continue
            ql = [(q[k] >> (8 * s)) & 255 for k in range(3)]
            qh = [(q[3 + k] >> (8 * s)) & 255 for k in range(3)]
            lo = [org[k] + ql[k] * scale[k] for k in range(3)]
            hi = [org[k] + qh[k] * scale[k] for k in range(3)]
            yield i, s, np.array(lo), np.array(hi), m


def cwbvh_subtree_tris(nodes, order, i):
    """Triangle ids (original) below CWBVH node i."""
    u = nodes.view(np.uint32)
    out = []
    imask = int(u[i, 3]) >> 24
    for (n, s, dl, dh, m) in _decode_cwbvh_child_boxes(nodes[i:i + 1]):
        if m & 0x80:
            first = int(u[i, 5]) + (m & 31)
            out += order[first:first + ((m >> 5) & 3) + 1].tolist()
        else:
            child = int(u[i, 4]) + bin(imask & ((1 << s) - 1)).count("1")
            out += cwbvh_subtree_tris(nodes, order, child)
    return out


def test_cwbvh_conservative_and_complete():
    v, t = random_soup(2000, 5)
    b = po.build(v, t, 8)
    nodes, order = b["nodes"], b["order"]
    assert sorted(order.tolist()) == list(range(len(t)))
    assert sorted(cwbvh_subtree_tris(nodes, order, 0)) == list(range(len(t)))   # every triangle reachable once
    assert b["max_leaf"] <= 3
    lo, hi = tri_boxes(v, t)
    u = nodes.view(np.uint32)
    # every slot's dequantised box contains the exact box of everything below it
    for (n, s, dl, dh, m) in _decode_cwbvh_child_boxes(nodes):
        if m & 0x80:
            first = int(u[n, 5]) + (m & 31)
            ids = order[first:first + ((m >> 5) & 3) + 1]
        else:
            imask = int(u[n, 3]) >> 24
            ids = cwbvh_subtree_tris(nodes, order, int(u[n, 4]) + bin(imask & ((1 << s) - 1)).count("1"))
        assert np.all(dl <= lo[ids].min(0)) and np.all(dh >= hi[ids].max(0))


def _rays(n, seed, lo=-0.5, hi=1.5):
    rng = np.random.default_rng(seed)
    o = rng.uniform(lo, hi, (n, 3)).astype(np.float32)
    d = rng.normal(size=(n, 3)).astype(np.float32)
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    return np.concatenate([o, np.zeros((n, 1), np.float32), d, np.full((n, 1), 1e30, np.float32)], 1)


def test_bvh4_layout_invariants():
    """4-wide nodes: child boxes are exact subtree boxes, every triangle sits in exactly one leaf, children come after
    their parent (pre-order), each node has >= 2 children unless the root is a wrapped leaf, and internal children
    were opened by largest area (a node keeps a leaf child only if no internal child could be opened)."""
    v, t = random_soup(1200, 13)
    b = po.build(v, t, 4)
    nodes = b["nodes"].reshape(-1, 32)
    ints = nodes.view(np.int32)
    order = b["order"]
    lo, hi = tri_boxes(v, t)
    seen = []

    def walk(j):
        tris = []
        for c in range(4):
            ci, k = ints[j, 24 + c], ints[j, 28 + c]
            if k == 0 and ci == -1:
                assert (nodes[j, [c, 4 + c, 8 + c, 12 + c, 16 + c, 20 + c]] == np.float32(3e37)).all()
                continue
            if k > 0:
                ids = order[ci:ci + k].tolist()
                seen.extend(ids)
            else:
                assert ci > j
                ids = walk(ci)
            blo = nodes[j, [c, 8 + c, 16 + c]]; bhi = nodes[j, [4 + c, 12 + c, 20 + c]]
            np.testing.assert_array_equal(blo, lo[ids].min(0)); np.testing.assert_array_equal(bhi, hi[ids].max(0))
            tris += ids
        return tris

    walk(0)
    assert sorted(seen) == list(range(len(t)))


@pytest.mark.parametrize("variant", [2, 4, 8])
@pytest.mark.parametrize("tri_test", [0, 1, 2])
def test_traversal_equals_brute_force(variant, tri_test):
    v, t = random_soup(1500, 7)
    b = po.build(v, t, variant)
    sc = po.Scene(variant, tri_test, b["nodes"], b["woop"], b["corners"], np.ones((1, 4), np.float32))
    rays = _rays(6000, 11)
    h1, h2 = sc.trace(rays), sc.trace_brute(rays)
    ids1, ids2 = h1[:, 3].view(np.int32), h2[:, 3].view(np.int32)
    assert (ids1 >= 0).sum() > 1000
    np.testing.assert_array_equal(ids1, ids2)
    np.testing.assert_array_equal(h1[:, :3], h2[:, :3])
    hany = sc.trace(rays, any_hit=True)
    np.testing.assert_array_equal(hany[:, 3].view(np.int32) >= 0, ids1 >= 0)


def test_watertight_triangle_test_has_no_gaps_at_shared_vertices():
    v, t = grid_mesh(17, seed=1, amp=0.0)   # flat grid: rays through vertices and edge midpoints must hit
    b = po.build(v, t, 8)
    rng = np.random.default_rng(5)
    targets = np.concatenate([v, 0.5 * (v[t[:, 0]] + v[t[:, 1]])])
    inner = np.all((targets[:, :2] > 0.01) & (targets[:, :2] < 0.99), axis=1)
    targets = targets[inner]
    o = targets + np.array([0.0, 0.0, 1.0], np.float32) + rng.uniform(-0.3, 0.3, targets.shape).astype(np.float32) * [1, 1, 0]
    d = targets - o
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    rays = np.concatenate([o, np.zeros((len(o), 1), np.float32), d, np.full((len(o), 1), 1e30, np.float32)], 1).astype(np.float32)
    # the triangle test alone (every triangle tested): watertight leaves no gap at shared vertices and edges
    sc = po.Scene(8, 2, b["nodes"], b["woop"], b["corners"], np.ones((1, 4), np.float32))
    h = sc.trace_brute(rays)
    assert np.all(h[:, 3].view(np.int32) >= 0)


# ------------------------------------------------------------------------------------------ legacy shading closed forms
class P(po.ctypes.Structure):
    _fields_ = [("pos", po.ctypes.c_float * 3), ("ax", po.ctypes.c_float * 3), ("ay", po.ctypes.c_float * 3),
                ("az", po.ctypes.c_float * 3), ("fx", po.ctypes.c_float), ("fy", po.ctypes.c_float),
                ("px", po.ctypes.c_float), ("py", po.ctypes.c_float), ("w", po.ctypes.c_int), ("h", po.ctypes.c_int),
                ("lens_r", po.ctypes.c_float), ("focus_d", po.ctypes.c_float), ("sun_to", po.ctypes.c_float * 3),
                ("sun_e", po.ctypes.c_float * 3), ("sky", po.ctypes.c_float * 3), ("sun_cos", po.ctypes.c_float),
                ("max_bounce", po.ctypes.c_int), ("seed", po.ctypes.c_uint32), ("eps", po.ctypes.c_float),
                ("near_t", po.ctypes.c_float), ("has_normals", po.ctypes.c_int), ("filter_bh", po.ctypes.c_int),
                ("ftab", po.ctypes.c_float * 65)]


def plane_params(w=32, h=32, sun=0.0, sky=(0.4, 0.5, 0.6)):
    p = P()
    # camera at height 1 looking down 45 degrees: the upper image rows see the sky
    pos = np.array([0, -2, 1.0]); tgt = np.array([0, 0, 0.0])
    f = tgt - pos; f /= np.linalg.norm(f); r = np.cross(f, [0, 0, 1.0]); r /= np.linalg.norm(r); u = np.cross(r, f)
    p.pos[:] = pos; p.ax[:] = r; p.ay[:] = u; p.az[:] = -f
    p.fx = p.fy = w * 0.5 / math.tan(math.radians(35)); p.px, p.py = w * 0.5, h * 0.5; p.w, p.h = w, h
    s = np.array([0.3, -0.4, 0.866]); s /= np.linalg.norm(s)
    p.sun_to[:] = s; p.sun_e[:] = (sun, sun, sun); p.sky[:] = sky; p.sun_cos = math.cos(math.radians(0.25))
    # ray offset 1e-4 x scene extent (the reference's rule; the plane spans 2000 units)
    p.max_bounce = 4; p.seed = 1000; p.eps = 1e-4 * 2000; p.near_t = 0.0; p.has_normals = 0; p.filter_bh = 0
    return p


def plane_scene(albedo=0.6, tri_test=1):
    v = np.array([[-1e3, -1e3, 0], [1e3, -1e3, 0], [1e3, 1e3, 0], [-1e3, 1e3, 0]], np.float32)
    t = np.array([[0, 1, 2], [0, 2, 3]], np.int32)
    b = po.build(v, t, 8)
    return po.Scene(8, tri_test, b["nodes"], b["woop"], b["corners"], np.array([[albedo] * 3 + [0]], np.float32))


@pytest.mark.parametrize("tri_test", [0, 1, 2])
def test_plane_under_sky_is_albedo_times_sky(tri_test):
    a = 0.6
    sc = plane_scene(a, tri_test)
    p = plane_params()
    acc, st = sc.render(p, 8, batch=8)
    img = acc[..., :3] / 8
    # every sample that sees the plane contributes exactly a * sky (its continuation always escapes), every other
    # sample exactly sky: a pixel's mean is sky * (1 - f (1 - a)) with f a multiple of 1/8 (horizon pixels mix)
    sky = np.array(p.sky[:], np.float32)
    f = (1.0 - img / sky) / (1.0 - a)
    np.testing.assert_allclose(f * 8, np.round(f * 8), atol=1e-4)
    np.testing.assert_allclose(f[..., 0], f[..., 1], atol=1e-5)
    plane = np.isclose(f[..., 0], 1.0, atol=1e-5)
    assert plane.sum() > 0.3 * plane.size and np.isclose(f[..., 0], 0.0, atol=1e-5).sum() > 0.05 * plane.size
    hits = np.round(f[..., 0] * 8).sum()
    assert st[0] == 8 * p.w * p.h + hits    # camera rays + one escaping continuation per plane sample


def test_plane_with_sun_closed_form():
    a, E = 0.5, 3.0
    sc = plane_scene(a)
    p = plane_params(sun=E, sky=(0.0, 0.0, 0.0))
    acc, st = sc.render(p, 16, batch=16)
    img = acc[..., 0] / 16
    cos_sun = float(np.array(p.sun_to[:])[2])
    expect = a * E * cos_sun / math.pi
    # each sample on the plane gives a E cos / pi (the sun cone spreads cos by < 0.5%), each sky sample 0; horizon
    # pixels mix, so a pixel is k/16 of the closed form for an integer k
    f = img / expect
    np.testing.assert_allclose(f * 16, np.round(f * 16), atol=16 * 5e-3)
    full = np.round(f * 16) == 16
    assert full.sum() > 0.3 * img.size
    np.testing.assert_allclose(img[full], expect, rtol=5e-3)


# ---------------------------------------------------------------------------------------------- H-PLOC restatement
def test_expand_bits21_closed_form():
    assert po.expand_bits21(0x1fffff) == 0x1249249249249249
    assert po.expand_bits21(0) == 0
    assert po.expand_bits21(1) == 1 and po.expand_bits21(2) == 8 and po.expand_bits21(0x100000) == 1 << 60
    rng = np.random.default_rng(3)
    for v in rng.integers(0, 1 << 21, 50):
        want = 0
        for b in range(21):
            want |= ((int(v) >> b) & 1) << (3 * b)
        assert po.expand_bits21(int(v)) == want


def _interleave(x, y, z, mb):
    m = 0
    for b in range(mb):
        m |= ((x >> b) & 1) << (3 * b + 2) | ((y >> b) & 1) << (3 * b + 1) | ((z >> b) & 1) << (3 * b)
    return m


def test_hploc_keys_against_direct_interleave():
    rng = np.random.default_rng(5)
    for n in (2, 7, 1000, 70000):
        c = rng.random((n, 3)).astype(np.float32) * np.float32(3) - np.float32(1)
        lo, hi = c - np.float32(0.01), c + np.float32(0.02)
        box, keys, mb, ib = po.hp_keys(lo, hi)
        assert ib == max(1, int(np.ceil(np.log2(n)))) and mb == min(21, (64 - ib) // 3)
        cen = (lo + hi) * np.float32(0.5)
        np.testing.assert_array_equal(box[:3], cen.min(0)); np.testing.assert_array_equal(box[3:], cen.max(0))
        s = np.float32(1 << mb)
        for i in rng.integers(0, n, 40):
            q = []
            for a in range(3):
                e = np.float32(box[3 + a] - box[a])
                t = np.float32(np.float32(cen[i, a] - box[a]) / e)
                q.append(int(min(max(t * s, np.float32(0)), s - np.float32(1))))
            assert int(keys[i]) == (_interleave(q[0], q[1], q[2], mb) << ib) | int(i)
        assert len(np.unique(keys)) == n


def test_hploc_30bit_keys_closed_form():
    """30-bit mode: 10 bits per axis whatever n is; the corners of the centroid box map to code 0 and 2^30 - 1, a
    centroid at 0.5 of the box on every axis to the code of (512, 512, 512), and the key order is (code, index)."""
    rng = np.random.default_rng(9)
    for n in (3, 1000, 70000):
        c = rng.random((n, 3)).astype(np.float32)
        c[0] = 0.0; c[1] = 1.0; c[2] = 0.5   # box corners and its centre (the other centroids lie inside [0, 1))
        lo, hi = c, c
        box, keys, mb, ib = po.hp_keys(lo, hi, 30)
        assert mb == 10 and ib == max(1, int(np.ceil(np.log2(n))))
        code = keys >> np.uint64(ib)
        assert int(code[0]) == 0 and int(code[1]) == (1 << 30) - 1
        assert int(code[2]) == 7 << 27   # 512 = bit 9 on each axis: bits 29, 28, 27
        assert int(code.max()) < 1 << 30
        np.testing.assert_array_equal(keys & np.uint64((1 << ib) - 1), np.arange(n, dtype=np.uint64))
        # the adaptive mode refines the same order: its codes shifted down to 10 bits per axis are the 30-bit codes
        _, ka, mba, iba = po.hp_keys(lo, hi, 0)
        assert mba > 10
        np.testing.assert_array_equal((ka >> np.uint64(iba)) >> np.uint64(3 * (mba - 10)), code)
        for i in rng.integers(0, n, 50):
            q = [int(min(max(np.float32(c[i, a] - box[a]) / np.float32(box[3 + a] - box[a]) * np.float32(1024), 0), 1023))
                 for a in range(3)]
            assert int(code[i]) == _interleave(q[0], q[1], q[2], 10)


def _hp_invariants(lo, hi, nodes, root, n):
    """2n - 1 nodes, every primitive once, parent box = union of children, children before parents."""
    ints = nodes.view(np.int32)
    assert root == (2 * n - 2 if n > 1 else 0)
    seen = sorted(ints[:n, 3].tolist())
    assert seen == list(range(n)) and (ints[:n, 7] == -1).all()
    np.testing.assert_array_equal(nodes[:n, :3], lo[ints[:n, 3]]); np.testing.assert_array_equal(nodes[:n, 4:7], hi[ints[:n, 3]])
    kids = ints[n:, [3, 7]]
    assert (kids < np.arange(n, 2 * n - 1)[:, None]).all()
    assert sorted(kids.ravel().tolist()) == list(range(2 * n - 2))   # every node but the root has one parent
    np.testing.assert_array_equal(nodes[n:, :3], np.minimum(nodes[kids[:, 0], :3], nodes[kids[:, 1], :3]))
    np.testing.assert_array_equal(nodes[n:, 4:7], np.maximum(nodes[kids[:, 0], 4:7], nodes[kids[:, 1], 4:7]))


@pytest.mark.parametrize("n", [1, 2, 3, 16, 17, 32, 33, 500, 20000])
def test_hploc_tree_invariants(n):
    rng = np.random.default_rng(n)
    c = rng.random((n, 3)).astype(np.float32)
    lo, hi = c - np.float32(0.005), c + rng.random((n, 3)).astype(np.float32) * np.float32(0.02)
    nodes, root, keys = po.hp_build(lo, hi)
    assert (np.diff(keys.astype(np.uint64)) > 0).all()
    _hp_invariants(lo, hi, nodes, root, n)


def _half_area(lo, hi):
    d = (hi - lo).astype(np.float32)
    return np.float32(np.float32(d[0] * d[1]) + np.float32(d[1] * d[2])) + np.float32(d[2] * d[0])


@pytest.mark.parametrize("n", [5, 9, 16, 17])
def test_hploc_small_trees_are_sequences_of_mutual_nearest_merges(n):
    """n <= 17: no LBVH node below the root holds more than 16 primitives, so the whole build is one PLOC window at
    the root over the leaves in Morton order. Replaying it from the Morton-ordered leaves, every
    internal node (in id order) must join two current clusters that are each other's nearest neighbour within 8
    positions by merged surface area (ties by position), and the survivors keep their order."""
    rng = np.random.default_rng(100 + n)
    c = rng.random((n, 3)).astype(np.float32)
    lo, hi = c - np.float32(0.01), c + np.float32(0.01)
    nodes, root, _ = po.hp_build(lo, hi)
    ints = nodes.view(np.int32)
    clusters = list(range(n))   # node ids in lane order
    box = {i: (nodes[i, :3].copy(), nodes[i, 4:7].copy()) for i in range(2 * n - 1)}

    def best_pairs(cl):
        m = len(cl)
        best = [None] * m
        for a in range(m):
            for r in range(1, 9):
                b = a + r
                if b >= m:
                    break
                l = np.minimum(box[cl[a]][0], box[cl[b]][0]); h = np.maximum(box[cl[a]][1], box[cl[b]][1])
                key = (int(_half_area(l, h).view(np.uint32)), a, b)
                for x in (a, b):
                    if best[x] is None or key < best[x]:
                        best[x] = key
        return best

    nid = n
    while len(clusters) > 1:
        best = best_pairs(clusters)
        nn = [bk[2] if bk[1] == k else bk[1] for k, bk in enumerate(best)]
        merged, keep = {}, []
        for k in range(len(clusters)):
            p = nn[k]
            if nn[p] == k and k < p:
                assert ints[nid, 3] == clusters[k] and ints[nid, 7] == clusters[p], (nid, k, p)
                merged[k] = nid
                nid += 1
        for k in range(len(clusters)):
            p = nn[k]
            if nn[p] == k and k > p:
                continue
            keep.append(merged.get(k, clusters[k]))
        assert merged, "an iteration must merge at least one pair"
        clusters = keep
    assert clusters == [root]


def test_tree_from_hploc_sah_collapse_against_brute_force():
    rng = np.random.default_rng(8)
    n = 300
    c = rng.random((n, 3)).astype(np.float32)
    lo, hi = c - np.float32(0.01), c + np.float32(0.03)
    nodes, root, _ = po.hp_build(lo, hi)
    ints = nodes.view(np.int32)

    def area(h):
        d = nodes[h, 4:7] - nodes[h, :3]
        return 2.0 * float(np.float32(np.float32(d[0] * d[1]) + np.float32(d[1] * d[2])) + np.float32(d[2] * d[0]))

    memo = {}

    def best(h):   # (cost, leaf?, count) by recursion from the definition
        if h in memo:
            return memo[h]
        if h < n:
            r = (area(h), True, 1)
        else:
            a, b = best(ints[h, 3]), best(ints[h, 7])
            cnt = a[2] + b[2]
            inner = area(h) + a[0] + b[0]
            leaf = area(h) * cnt if cnt <= 4 else 1e300
            r = (min(leaf, inner), leaf <= inner, cnt)
        memo[h] = r
        return r

    import sys
    sys.setrecursionlimit(10000)
    t = po.tree_from_hploc(nodes, n, root, 4)
    tree = t["tree"]
    # every output leaf is a subtree that the recursion marks as leaf, and the leaves cover every primitive once
    assert sorted(t["prim"].tolist()) == list(range(n))
    leaves = tree[tree[:, 3] > 0]
    assert (leaves[:, 3] <= 4).all()
    cost_tree = 0.0
    for k in range(len(tree)):
        lo_, hi_ = t["tree_box"][k, :3], t["tree_box"][k, 3:]
        d = hi_ - lo_
        A = 2.0 * float(np.float32(np.float32(d[0] * d[1]) + np.float32(d[1] * d[2])) + np.float32(d[2] * d[0]))
        cost_tree += A * tree[k, 3] if tree[k, 3] > 0 else A
    assert cost_tree == pytest.approx(best(root)[0], rel=1e-12)
