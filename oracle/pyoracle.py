"""ctypes wrapper of the oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: the CPU parity checker. Imported by the tests, the smoke check and bench.py's CPU-baseline
leg; never by the product.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

_P = ctypes.c_void_p
_FP = ctypes.POINTER(ctypes.c_float)
_IP = ctypes.POINTER(ctypes.c_int)
_DP = ctypes.POINTER(ctypes.c_double)


class OrcTree(ctypes.Structure):
    _fields_ = [("num_nodes", ctypes.c_int), ("num_prims", ctypes.c_int), ("max_depth", ctypes.c_int),
                ("node", _IP), ("box", _FP), ("prim", _IP)]


class OrcLayout(ctypes.Structure):
    _fields_ = [("num_float4", ctypes.c_int), ("nodes", _FP), ("num_prims", ctypes.c_int), ("order", _IP),
                ("depth", ctypes.c_int), ("max_leaf", ctypes.c_int), ("sah", ctypes.c_double)]


class OrcScene(ctypes.Structure):
    _fields_ = [("variant", ctypes.c_int), ("tri_test", ctypes.c_int), ("nodes", _FP), ("root", ctypes.c_int),
                ("woop", _FP), ("corners", _FP), ("normals", _FP), ("albedo", _FP), ("num_tris", ctypes.c_int),
                ("tlas", _FP), ("tlas_root", ctypes.c_int), ("inst", _FP), ("num_inst", ctypes.c_int),
                ("full", ctypes.c_int), ("mats", _FP), ("num_mats", ctypes.c_int), ("tri_light_pdf", _FP), ("em_tri", _IP),
                ("em_cdf", _FP), ("num_em", ctypes.c_int), ("env_tex", _FP), ("env_marg", _FP), ("env_cond", _FP),
                ("env_w", ctypes.c_int), ("env_h", ctypes.c_int), ("env_rot", ctypes.c_float), ("p_sun", ctypes.c_float),
                ("p_tri", ctypes.c_float), ("p_env", ctypes.c_float), ("light_mode", ctypes.c_int), ("rr", ctypes.c_int),
                ("rr_depth", ctypes.c_int), ("gd_energy", ctypes.c_int), ("env_uniform", ctypes.c_int),
                ("em_tris", _FP), ("mat_light_pdf", _FP), ("maps", _IP), ("num_maps", ctypes.c_int), ("uvs", _FP),
                ("tans", _FP), ("tex_images", _IP), ("texels", ctypes.POINTER(ctypes.c_uint32)), ("srgb", _FP),
                ("tex_nearest", ctypes.c_int), ("bump_shadow", ctypes.c_int)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("oracle not built: python oracle/build.py")
        L = ctypes.CDLL(LIB_PATH)
        L.orc_hash32.argtypes = [ctypes.c_uint32]; L.orc_hash32.restype = ctypes.c_uint32
        L.orc_rng4.argtypes = [ctypes.c_uint32] * 5; L.orc_rng4.restype = ctypes.c_float
        L.orc_blackman_harris_table.argtypes = [_FP, ctypes.c_int, ctypes.c_double]
        L.orc_cosine_hemisphere.argtypes = [_FP, ctypes.c_float, ctypes.c_float, _FP]
        L.orc_uniform_cone.argtypes = [_FP, ctypes.c_float, ctypes.c_float, ctypes.c_float, _FP]
        L.orc_build_sah.argtypes = [_FP, ctypes.c_int, _IP, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.POINTER(OrcTree)]
        L.orc_build_sah.restype = ctypes.c_int
        L.orc_tree_free.argtypes = [ctypes.POINTER(OrcTree)]
        L.orc_layout_bvh2.argtypes = [ctypes.POINTER(OrcTree), ctypes.POINTER(OrcLayout)]
        L.orc_collapse_cwbvh.argtypes = [ctypes.POINTER(OrcTree), ctypes.POINTER(OrcLayout)]
        L.orc_layout_bvh4.argtypes = [ctypes.POINTER(OrcTree), ctypes.POINTER(OrcLayout)]
        L.orc_layout_free.argtypes = [ctypes.POINTER(OrcLayout)]
        L.orc_pack_woop.argtypes = [_FP, _IP, _IP, ctypes.c_int, _FP]
        L.orc_pack_corners.argtypes = [_FP, _IP, _IP, _IP, ctypes.c_int, _FP]
        L.orc_trace.argtypes = [ctypes.POINTER(OrcScene), _FP, ctypes.c_int, ctypes.c_int, _FP, ctypes.c_int]
        L.orc_trace_brute.argtypes = [ctypes.POINTER(OrcScene), _FP, ctypes.c_int, _FP, ctypes.c_int]
        L.orc_trace_work.argtypes = [ctypes.POINTER(OrcScene), _FP, ctypes.c_int, _DP, _DP]
        L.orc_render.argtypes = [ctypes.POINTER(OrcScene), _P, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                 ctypes.c_int, ctypes.c_int, _FP, _DP]
        L.orc_render.restype = ctypes.c_int
        L.orc_camera_ray.argtypes = [_P, ctypes.c_int, ctypes.c_uint32, _FP, _FP]
        _U64P = ctypes.POINTER(ctypes.c_uint64)
        L.orc_hp_centroid_box.argtypes = [_FP, _FP, ctypes.c_int, _FP]
        L.orc_trace_inst.argtypes = [ctypes.POINTER(OrcScene), _FP, ctypes.c_int, ctypes.c_int, _FP, _IP, ctypes.c_int]
        L.orc_inst_boxes.argtypes = [_FP, ctypes.c_int, _FP, _FP, _FP]
        L.orc_expand_bits21.argtypes = [ctypes.c_uint32]; L.orc_expand_bits21.restype = ctypes.c_uint64
        L.orc_hp_key_bits.argtypes = [ctypes.c_int, _IP, _IP]
        L.orc_hp_keys.argtypes = [_FP, _FP, ctypes.c_int, _FP, ctypes.c_int, ctypes.c_int, _U64P]
        L.orc_hp_build.argtypes = [_FP, _FP, ctypes.c_int, _FP, _U64P]
        L.orc_hp_build_mode.argtypes = [_FP, _FP, ctypes.c_int, ctypes.c_int, _FP, _U64P]
        L.orc_hp_key_bits_mode.argtypes = [ctypes.c_int, ctypes.c_int, _IP, _IP]
        L.orc_hp_to_bvh2.argtypes = [_FP, ctypes.c_int, _FP]
        L.orc_tree_from_hploc.argtypes = [_FP, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.POINTER(OrcTree)]
        _F = ctypes.c_float
        L.orc_bsdf_eval.argtypes = [_FP, _FP, _FP, _F, ctypes.c_int, _F, _F, _FP, _FP]
        L.orc_texture_sample.argtypes = [_IP, ctypes.POINTER(ctypes.c_uint32), _FP, _F, _F, ctypes.c_int, _FP]
        L.orc_bsdf_sample.argtypes = [_FP, _FP, _F, ctypes.c_int, _F, _F, _F, _F, _F, _FP, _FP, _FP]
        L.orc_bsdf_sample.restype = ctypes.c_int
        L.orc_bsdf_sample_batch.argtypes = [_FP, _FP, _F, ctypes.c_int, ctypes.c_int, ctypes.c_uint64, _FP, _FP, _FP, _IP]
        L.orc_bsdf_pdf_batch.argtypes = [_FP, _FP, _F, ctypes.c_int, _FP, ctypes.c_int, _FP]
        L.orc_fresnel_dielectric.argtypes = [_F, _F]; L.orc_fresnel_dielectric.restype = _F
        L.orc_fresnel_f82.argtypes = [_FP, _FP, _F, _FP]
        L.orc_ggx_d.argtypes = [_F, _F]; L.orc_ggx_d.restype = _F
        L.orc_ggx_g2.argtypes = [_F, _F, _F]; L.orc_ggx_g2.restype = _F
        L.orc_quad_albedo.argtypes = [ctypes.c_int, _F, _F, _FP]; L.orc_quad_albedo.restype = ctypes.c_double
        L.orc_env_build.argtypes = [_FP, ctypes.c_int, ctypes.c_int, _F, _FP, _FP, _FP]
        L.orc_env_pdf.argtypes = [ctypes.POINTER(OrcScene), _FP]; L.orc_env_pdf.restype = _F
        L.orc_env_sample_batch.argtypes = [ctypes.POINTER(OrcScene), ctypes.c_int, ctypes.c_uint64, _FP, _FP]
        L.orc_emitter_tables.argtypes = [_FP, _IP, ctypes.c_int, _FP, _IP, _FP, _FP]
        L.orc_emitter_tables.restype = ctypes.c_int
        L.orc_inst_emitters.argtypes = [_FP, ctypes.c_int, _IP, _FP, _IP, _FP, ctypes.c_int, _FP, _FP, _FP]
        L.orc_inst_emitters.restype = ctypes.c_int
        L.orc_display.argtypes = [_FP, ctypes.c_int, ctypes.c_int, _F, ctypes.c_int, ctypes.POINTER(ctypes.c_uint8)]
        L.orc_srgb_encode8.argtypes = [_F]; L.orc_srgb_encode8.restype = ctypes.c_int
        L.orc_srgb_steps_check.argtypes = [_FP, _F, _F, ctypes.c_int]; L.orc_srgb_steps_check.restype = ctypes.c_uint64
        _lib = L
    return _lib


def _f(a):
    return a.ctypes.data_as(_FP)


def _i(a):
    return a.ctypes.data_as(_IP)


def hash32(x):
    return lib().orc_hash32(x)


def rng4(pix, smp, depth, dim, seed):
    return lib().orc_rng4(pix, smp, depth, dim, seed)


def blackman_harris_table(n=64, width=1.5):
    t = np.zeros(n + 1, dtype=np.float32)
    lib().orc_blackman_harris_table(_f(t), n, width)
    return t


def cosine_hemisphere(n, u0, u1):
    a = np.asarray(n, dtype=np.float32); o = np.zeros(3, dtype=np.float32)
    lib().orc_cosine_hemisphere(_f(a), u0, u1, _f(o))
    return o


def uniform_cone(axis, cos_max, u0, u1):
    a = np.asarray(axis, dtype=np.float32); o = np.zeros(3, dtype=np.float32)
    lib().orc_uniform_cone(_f(a), cos_max, u0, u1, _f(o))
    return o


def build_sah(vertices, triangles, bins=16, leaf_max=4):
    v = np.ascontiguousarray(vertices, dtype=np.float32).reshape(-1, 3)
    t = np.ascontiguousarray(triangles, dtype=np.int32).reshape(-1, 3)
    tr = OrcTree()
    if lib().orc_build_sah(_f(v), v.shape[0], _i(t), t.shape[0], bins, leaf_max, ctypes.byref(tr)) != 0:
        raise ValueError("orc_build_sah: bad input")
    return tr


def tree_arrays(tr):
    n = tr.num_nodes
    node = np.ctypeslib.as_array(tr.node, shape=(n * 4,)).reshape(n, 4).copy()
    box = np.ctypeslib.as_array(tr.box, shape=(n * 6,)).reshape(n, 6).copy()
    prim = np.ctypeslib.as_array(tr.prim, shape=(tr.num_prims,)).copy()
    return dict(tree=node, tree_box=box, prim=prim, depth=tr.max_depth)


def _box4(a):
    a = np.ascontiguousarray(a, dtype=np.float32)
    if a.ndim == 2 and a.shape[1] == 3:
        a = np.concatenate([a, np.zeros((len(a), 1), np.float32)], 1)
    return np.ascontiguousarray(a.reshape(-1, 4))


def hp_key_bits(n, morton_bits=0):
    """(bits per axis, index bits) of the keys; morton_bits 0 = adaptive (the reference), 30 = 10 bits per axis"""
    mb, ib = ctypes.c_int(), ctypes.c_int()
    lib().orc_hp_key_bits_mode(n, morton_bits, ctypes.byref(mb), ctypes.byref(ib))
    return mb.value, ib.value


def hp_keys(lo, hi, morton_bits=0):
    """Centroid box and unsorted 64-bit keys of primitive boxes (lo/hi: (n, 3) or (n, 4))."""
    lo, hi = _box4(lo), _box4(hi)
    n = len(lo)
    box = np.zeros(6, np.float32)
    lib().orc_hp_centroid_box(_f(lo), _f(hi), n, _f(box))
    mb, ib = hp_key_bits(n, morton_bits)
    keys = np.zeros(n, np.uint64)
    lib().orc_hp_keys(_f(lo), _f(hi), n, _f(box), mb, ib, keys.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)))
    return box, keys, mb, ib


def expand_bits21(v):
    return lib().orc_expand_bits21(v)


def hp_build(lo, hi, morton_bits=0):
    """H-PLOC build: returns (nodes (2n-1, 8) float32, root, sorted keys)."""
    lo, hi = _box4(lo), _box4(hi)
    n = len(lo)
    nodes = np.zeros((max(2 * n - 1, 1), 8), np.float32)
    keys = np.zeros(n, np.uint64)
    root = lib().orc_hp_build_mode(_f(lo), _f(hi), n, morton_bits, _f(nodes),
                                   keys.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)))
    return nodes, root, keys


def hp_to_bvh2(nodes, n):
    h = np.ascontiguousarray(nodes, dtype=np.float32)
    out = np.zeros((max(n - 1, 1), 16), np.float32)
    lib().orc_hp_to_bvh2(_f(h), n, _f(out))
    return out


def tree_from_hploc(nodes, n, root, leaf_max):
    h = np.ascontiguousarray(nodes, dtype=np.float32)
    tr = OrcTree()
    lib().orc_tree_from_hploc(_f(h), n, root, leaf_max, ctypes.byref(tr))
    out = tree_arrays(tr)
    lib().orc_tree_free(ctypes.byref(tr))
    return out


def hp_canonical(nodes, n, root):
    """Canonical pre-order form of an H-PLOC tree, independent of internal node numbering: a list of tuples
    (lo/hi bits) + ('leaf', primitive) or ('node', left index, right index) with pre-order indices."""
    h = np.ascontiguousarray(nodes, dtype=np.float32)
    bits = h.view(np.uint32)
    out, st = [], [(root, -1, 0)]
    while st:
        x, parent, side = st.pop()
        me = len(out)
        if parent >= 0:
            out[parent][2 + side] = me
        if x < n:
            out.append([tuple(bits[x, [0, 1, 2, 4, 5, 6]].tolist()), "leaf", int(bits[x, 3].view(np.int32)), -1])
        else:
            out.append([tuple(bits[x, [0, 1, 2, 4, 5, 6]].tolist()), "node", -1, -1])
            st.append((int(h[x, 7:8].view(np.int32)[0]), me, 1))
            st.append((int(h[x, 3:4].view(np.int32)[0]), me, 0))
    return [tuple(r) for r in out]


def _layout(tr, variant):
    L = OrcLayout()
    if variant == 2:
        lib().orc_layout_bvh2(ctypes.byref(tr), ctypes.byref(L))
    elif variant == 4:
        lib().orc_layout_bvh4(ctypes.byref(tr), ctypes.byref(L))
    else:
        lib().orc_collapse_cwbvh(ctypes.byref(tr), ctypes.byref(L))
    nodes = np.ctypeslib.as_array(L.nodes, shape=(L.num_float4 * 4,)).copy()
    order = np.ctypeslib.as_array(L.order, shape=(L.num_prims,)).copy()
    out = dict(nodes=nodes.reshape(-1, {2: 16, 4: 32, 8: 20}[variant]), order=order, layout_depth=L.depth,
               max_leaf=L.max_leaf, sah=L.sah)
    lib().orc_layout_free(ctypes.byref(L))
    return out


def tri_boxes(vertices, triangles):
    """Per-triangle boxes (n, 4) lo and hi of a mesh (exact min / max of the corners)."""
    v = np.ascontiguousarray(vertices, dtype=np.float32).reshape(-1, 3)
    t = np.ascontiguousarray(triangles, dtype=np.int32).reshape(-1, 3)
    c = v[t]
    z = np.zeros((len(t), 1), np.float32)
    return np.concatenate([c.min(1), z], 1), np.concatenate([c.max(1), z], 1)


def build(vertices, triangles, variant=8, bins=16, leaf_max=4, builder="sah", hnodes=None, root=None):
    """Full oracle build: tree (binned SAH, or the SAH leaf collapse of H-PLOC nodes -- the oracle's own H-PLOC build
    unless hnodes/root are given), leaf_max 1 for the CWBVH collapse, then layout, Woop rows, corners."""
    v = np.ascontiguousarray(vertices, dtype=np.float32).reshape(-1, 3)
    t = np.ascontiguousarray(triangles, dtype=np.int32).reshape(-1, 3)
    if builder == "hploc":
        if hnodes is None:
            hnodes, root, _ = hp_build(*tri_boxes(v, t))
        tr = OrcTree()
        h = np.ascontiguousarray(hnodes, dtype=np.float32)
        lib().orc_tree_from_hploc(_f(h), len(t), int(root), 1 if variant == 8 else leaf_max, ctypes.byref(tr))
    else:
        tr = build_sah(v, t, bins, 1 if variant == 8 else leaf_max)
    try:
        out = tree_arrays(tr)
        out.update(_layout(tr, variant))
    finally:
        lib().orc_tree_free(ctypes.byref(tr))
    n = len(out["order"])
    woop = np.zeros((n, 3, 4), dtype=np.float32)
    corners = np.zeros((n, 3, 4), dtype=np.float32)
    order = np.ascontiguousarray(out["order"], dtype=np.int32)
    lib().orc_pack_woop(_f(v), _i(t), _i(order), n, _f(woop))
    lib().orc_pack_corners(_f(v), _i(t), None, _i(order), n, _f(corners))
    out.update(woop=woop, corners=corners, variant=variant)
    return out


class Scene:
    """Arrays of a built scene in leaf order, kept alive for the C calls."""

    def __init__(self, variant, tri_test, nodes, woop, corners, albedo, normals=None, root=0):
        self.nodes = np.ascontiguousarray(nodes, dtype=np.float32).reshape(-1)
        self.woop = np.ascontiguousarray(woop, dtype=np.float32).reshape(-1)
        self.corners = np.ascontiguousarray(corners, dtype=np.float32).reshape(-1)
        self.albedo = np.ascontiguousarray(albedo, dtype=np.float32).reshape(-1)
        self.normals = None if normals is None or len(normals) == 0 else np.ascontiguousarray(normals, dtype=np.float32).reshape(-1)
        self.c = OrcScene(variant, tri_test, _f(self.nodes), root, _f(self.woop), _f(self.corners),
                          _f(self.normals) if self.normals is not None else None, _f(self.albedo), len(self.corners) // 12,
                          None, 0, None, 0)

    def set_full(self, mats, order, sun_on=1, light_mode=0, rr=0, rr_depth=3, gd_energy=1, env_uniform=0, env_tex=None,
                 env_w=0, env_h=0, env_rot_deg=0.0):
        """Full material integrator: materials (m, 16), the leaf order (leaf position -> mesh triangle) for the
        emitter CDF, options, and an optional environment (texels (h, w, 4) or (w*h, 4), radiance x strength). The
        emitter tables and the environment CDFs are built by the oracle; the light types present share the selection
        probability equally, the sun always, emitters and environment unless light_mode is 2 (BSDF only)."""
        self.mats = np.ascontiguousarray(mats, dtype=np.float32).reshape(-1, 16)
        nt = len(self.corners) // 12
        self.em_tri, self.em_cdf, self.tri_light_pdf = emitter_tables(self.corners.reshape(-1, 3, 4), order, self.mats)
        c = self.c
        c.full = 1; c.mats = _f(self.mats); c.num_mats = len(self.mats)
        c.num_em = len(self.em_tri)
        c.em_tri = _i(self.em_tri); c.em_cdf = _f(self.em_cdf); c.tri_light_pdf = _f(self.tri_light_pdf)
        if env_tex is not None and env_w > 0:
            self.env_tex, self.env_marg, self.env_cond = env_build(np.asarray(env_tex, np.float32).reshape(-1, 4)[:, :3].reshape(env_h, env_w, 3), 1.0)
            c.env_tex = _f(self.env_tex); c.env_marg = _f(self.env_marg); c.env_cond = _f(self.env_cond)
            c.env_w, c.env_h = env_w, env_h
            c.env_rot = np.float32(env_rot_deg) * np.float32(3.14159265) / np.float32(180.0)
        else:
            c.env_w = c.env_h = 0
        tri = c.num_em > 0 and light_mode != 2
        env = c.env_w > 0 and light_mode != 2
        nt_types = int(bool(sun_on)) + int(tri) + int(env)
        p = np.float32(1.0) / np.float32(nt_types) if nt_types else np.float32(0)
        c.p_sun = p if sun_on else 0.0
        c.p_tri = p if tri else 0.0
        c.p_env = p if env else 0.0
        c.light_mode, c.rr, c.rr_depth, c.gd_energy, c.env_uniform = light_mode, rr, rr_depth, gd_energy, env_uniform
        assert nt == len(self.tri_light_pdf)

    def set_textures(self, maps, uvs, tex_images, texels, tans=None, nearest=0, bump_shadow=0):
        """Texture maps of the full integrator: bindings (m, 8) int32, corner uvs in leaf order (n, 6) or (n, 2, 4),
        images (k, 4) int32 (first, w, h, srgb), texels uint32, tangents in leaf order (n, 3, 4) or None; the sRGB
        decode table is the oracle's own (srgb_table)."""
        self.maps = np.ascontiguousarray(maps, np.int32).reshape(-1, 8)
        u = np.asarray(uvs, np.float32)
        if u.ndim == 2 and u.shape[1] == 6:
            u = np.concatenate([u, np.zeros((len(u), 2), np.float32)], 1)
        self.uvs = np.ascontiguousarray(u, np.float32).reshape(-1)
        self.tex_images = np.ascontiguousarray(tex_images, np.int32).reshape(-1)
        self.texels = np.ascontiguousarray(texels, np.uint32).reshape(-1)
        self.srgb = srgb_table()
        self.tans = None if tans is None or len(tans) == 0 else np.ascontiguousarray(tans, np.float32).reshape(-1)
        c = self.c
        c.maps = _i(self.maps); c.num_maps = len(self.maps); c.uvs = _f(self.uvs)
        c.tans = _f(self.tans) if self.tans is not None else None
        c.tex_images = _i(self.tex_images)
        c.texels = self.texels.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32))
        c.srgb = _f(self.srgb)
        c.tex_nearest, c.bump_shadow = int(nearest), int(bump_shadow)

    def set_inst_emitters(self, mesh_tris, order):
        """Two-level full-integrator scene (after set_instances and set_full): the frame's world emitters, computed
        by the oracle from the instance records and the BLAS corners; returns (em_tris (n, 3, 4), em_cdf, mlp)."""
        L = lib()
        self.mesh_tris = np.ascontiguousarray(mesh_tris, np.int32).reshape(-1)
        self.order = np.ascontiguousarray(order, np.int32)
        args = (_f(self.inst), len(self.inst), _i(self.mesh_tris), _f(self.corners), _i(self.order), _f(self.mats), len(self.mats))
        n = L.orc_inst_emitters(*args, None, None, None)
        self.em_tris = np.zeros((max(n, 1), 3, 4), np.float32)
        self.em_cdf = np.zeros(n + 1, np.float32)
        self.mlp = np.zeros(len(self.mats), np.float32)
        L.orc_inst_emitters(*args, _f(self.em_tris), _f(self.em_cdf), _f(self.mlp))
        c = self.c
        c.em_tris = _f(self.em_tris); c.mat_light_pdf = _f(self.mlp); c.em_cdf = _f(self.em_cdf); c.num_em = n
        tri = n > 0 and c.light_mode != 2
        env = c.env_w > 0 and c.light_mode != 2
        k = int(c.p_sun > 0) + int(tri) + int(env)
        p = np.float32(1.0) / np.float32(k) if k else np.float32(0)
        c.p_tri = p if tri else 0.0
        c.p_env = p if env else 0.0
        c.p_sun = p if c.p_sun > 0 else 0.0
        return self.em_tris[:n], self.em_cdf, self.mlp

    def set_instances(self, inst, mesh_box):
        """Two-level scene: instance records (n, 28) float32 (world-to-object rows, object-to-world rows, root bits,
        mesh bits, det sign, pad) and object-space mesh boxes (m, 8) (lo, hi). Builds the TLAS with the oracle's H-PLOC
        restatement over the instance boxes; returns the TLAS BVH2 nodes."""
        self.inst = np.ascontiguousarray(inst, dtype=np.float32).reshape(-1, 28)
        mb = np.ascontiguousarray(mesh_box, dtype=np.float32).reshape(-1, 8)
        n = len(self.inst)
        lo, hi = np.zeros((n, 4), np.float32), np.zeros((n, 4), np.float32)
        lib().orc_inst_boxes(_f(self.inst), n, _f(mb), _f(lo), _f(hi))
        hn, root, _ = hp_build(lo, hi)
        self.tlas = hp_to_bvh2(hn, n)
        self.inst_boxes = (lo, hi)
        self.c.tlas = _f(self.tlas)
        self.c.tlas_root = n - 2 if n > 1 else 0
        self.c.inst = _f(self.inst)
        self.c.num_inst = n
        return self.tlas

    def trace_inst(self, rays, any_hit=False, threads=8):
        r = np.ascontiguousarray(rays, dtype=np.float32).reshape(-1, 8)
        h = np.zeros((r.shape[0], 4), dtype=np.float32)
        ids = np.zeros(r.shape[0], dtype=np.int32)
        lib().orc_trace_inst(ctypes.byref(self.c), _f(r), r.shape[0], 1 if any_hit else 0, _f(h), _i(ids), threads)
        return h, ids

    def trace(self, rays, any_hit=False, threads=8):
        r = np.ascontiguousarray(rays, dtype=np.float32).reshape(-1, 8)
        h = np.zeros((r.shape[0], 4), dtype=np.float32)
        lib().orc_trace(ctypes.byref(self.c), _f(r), r.shape[0], 1 if any_hit else 0, _f(h), threads)
        return h

    def trace_brute(self, rays, threads=8):
        r = np.ascontiguousarray(rays, dtype=np.float32).reshape(-1, 8)
        h = np.zeros((r.shape[0], 4), dtype=np.float32)
        lib().orc_trace_brute(ctypes.byref(self.c), _f(r), r.shape[0], _f(h), threads)
        return h

    def trace_work(self, rays):
        r = np.ascontiguousarray(rays, dtype=np.float32).reshape(-1, 8)
        nv, nt = ctypes.c_double(), ctypes.c_double()
        lib().orc_trace_work(ctypes.byref(self.c), _f(r), r.shape[0], ctypes.byref(nv), ctypes.byref(nt))
        return nv.value, nt.value

    def render(self, params, spp, sample_base=0, batch=None, bounces=None, threads=8, accum=None):
        """params: a ctypes structure with the cwp_params layout. Returns (accum (h, w, 4), stats[4])."""
        w, h = params.w, params.h
        clear = accum is None
        if accum is None:
            accum = np.zeros((h, w, 4), dtype=np.float32)
        stats = np.zeros(4, dtype=np.float64)
        batch = spp if batch is None else batch
        bounces = params.max_bounce if bounces is None else bounces
        rc = lib().orc_render(ctypes.byref(self.c), ctypes.cast(ctypes.byref(params), _P), spp, sample_base, batch, bounces,
                              threads, 1 if clear else 0, _f(accum), stats.ctypes.data_as(_DP))
        if rc:
            raise ValueError("orc_render: bad arguments")
        return accum, stats


def emitter_tables(corners, order, mats):
    """Emissive triangles by power in mesh order: (em_tri leaf positions, em_cdf, tri_light_pdf per leaf position)."""
    c = np.ascontiguousarray(corners, dtype=np.float32)
    o = np.ascontiguousarray(order, dtype=np.int32)
    m = np.ascontiguousarray(mats, dtype=np.float32)
    nt = len(o)
    et, ec, tp = np.zeros(max(nt, 1), np.int32), np.zeros(nt + 1, np.float32), np.zeros(nt, np.float32)
    n = lib().orc_emitter_tables(_f(c), _i(o), nt, _f(m), _i(et), _f(ec), _f(tp))
    return et[:n].copy(), ec[:n + 1].copy() if n else np.zeros(0, np.float32), tp


def env_build(rgb, strength):
    """Environment texels (w*h, 4), marginal (h + 1) and conditional (h, w + 1) CDFs of an (h, w, 3) image."""
    a = np.ascontiguousarray(rgb, dtype=np.float32)
    h, w = a.shape[:2]
    tex, marg, cond = np.zeros((w * h, 4), np.float32), np.zeros(h + 1, np.float32), np.zeros((h, w + 1), np.float32)
    lib().orc_env_build(_f(a), w, h, strength, _f(tex), _f(marg), _f(cond))
    return tex, marg, cond


def bsdf_eval(mat, wo, wi, eta, gd_energy=1, ux1=0.5, ux2=0.5):
    m = np.ascontiguousarray(mat, np.float32); a = np.asarray(wo, np.float32); b = np.asarray(wi, np.float32)
    f = np.zeros(3, np.float32); p = np.zeros(1, np.float32)
    lib().orc_bsdf_eval(_f(m), _f(a), _f(b), eta, gd_energy, ux1, ux2, _f(f), _f(p))
    return f, float(p[0])


def bsdf_sample_batch(mat, wo, eta, n, seed, gd_energy=1):
    m = np.ascontiguousarray(mat, np.float32); a = np.asarray(wo, np.float32)
    wi, w, p, ok = np.zeros((n, 3), np.float32), np.zeros((n, 3), np.float32), np.zeros(n, np.float32), np.zeros(n, np.int32)
    lib().orc_bsdf_sample_batch(_f(m), _f(a), eta, gd_energy, n, seed, _f(wi), _f(w), _f(p), _i(ok))
    return wi, w, p, ok


def bsdf_pdf_batch(mat, wo, eta, wi, gd_energy=1):
    m = np.ascontiguousarray(mat, np.float32); a = np.asarray(wo, np.float32)
    d = np.ascontiguousarray(wi, np.float32).reshape(-1, 3)
    p = np.zeros(len(d), np.float32)
    lib().orc_bsdf_pdf_batch(_f(m), _f(a), eta, gd_energy, _f(d), len(d), _f(p))
    return p


def material(base=(0.8, 0.8, 0.8), metallic=0.0, roughness=0.5, ior=1.5, specular_weight=1.0,
             specular_color=(1, 1, 1), transmission=0.0, emission=(0, 0, 0)):
    """16-float material record (OpenPBR defaults)."""
    return np.array(list(base) + [metallic] + list(specular_color) + [roughness] + list(emission) + [ior] +
                    [specular_weight, transmission, 0, 0], dtype=np.float32)


def camera_ray(params, pix, smp):
    o = np.zeros(3, dtype=np.float32); d = np.zeros(3, dtype=np.float32)
    lib().orc_camera_ray(ctypes.cast(ctypes.byref(params), _P), pix, smp, _f(o), _f(d))
    return o, d


def scene_from_setup(setup, tri_test=None, bins=16):
    """Oracle scene over a product setup's mesh, rebuilt by the oracle's own builder (not the product's arrays)."""
    import cwp
    v = setup.array(cwp.ARR_VERTICES)
    t = setup.array(cwp.ARR_TRIANGLES)
    mats = setup.array(cwp.ARR_MATERIAL)
    alb = setup.array(cwp.ARR_ALBEDO)
    info = setup.info
    b = build(v, t, info["variant"], bins, 4)
    corners = np.zeros((len(b["order"]), 3, 4), dtype=np.float32)
    order = np.ascontiguousarray(b["order"], dtype=np.int32)
    lib().orc_pack_corners(_f(np.ascontiguousarray(v, dtype=np.float32)), _i(np.ascontiguousarray(t, dtype=np.int32)),
                           _i(np.ascontiguousarray(mats, dtype=np.int32)), _i(order), len(order), _f(corners))
    b["corners"] = corners
    tt = info["tri_test"] if tri_test is None else tri_test
    nrm = None
    if setup.params().has_normals:   # smooth shading: the vertex normals at each corner, in the oracle's leaf order
        vn = setup.array(cwp.ARR_NORMALS)
        nrm = np.zeros((len(order), 3, 4), np.float32)
        nrm[..., :3] = vn[np.asarray(t, np.int64)[order]]
    sc = Scene(info["variant"], tt, b["nodes"], b["woop"], corners, alb, normals=nrm)
    if info.get("full"):
        env = setup.array(cwp.ARR_ENV_TEX) if info["env_w"] > 0 else None
        sc.set_full(setup.array(cwp.ARR_MATERIALS), order, info["sun_on"], info["light_mode"], info["rr"], info["rr_depth"],
                    info["gd_energy"], info["env_uniform"], env, info["env_w"], info["env_h"], info["env_rot_deg"])
    if info.get("textured"):
        # per-triangle corner uvs in mesh order, put into the oracle's leaf order; tangents computed by the oracle
        muv = setup.array(cwp.ARR_MESH_UVS).reshape(-1, 6)
        maps = setup.array(cwp.ARR_MAPS)
        tans = None
        if (maps[:, 2] >= 0).any():
            n = setup.array(cwp.ARR_NORMALS) if setup.params().has_normals else None
            tans = corner_tangents(v, n, t, muv)[order]
        sc.set_textures(maps, muv[order], setup.array(cwp.ARR_TEX_IMAGES), setup.array(cwp.ARR_TEXELS), tans,
                        info["tex_nearest"], info["bump_shadow"])
    return sc, b


def srgb_table():
    """sRGB decoding of the byte values (IEC 61966-2-1 EOTF in double precision, rounded to float32)."""
    c = np.arange(256, dtype=np.float64) / 255.0
    return np.where(c <= 0.04045, c / 12.92, ((c + 0.055) / 1.055) ** 2.4).astype(np.float32)


def corner_tangents(verts, normals, tris, uv):
    """Tangent frames per triangle corner (n, 3, 4) from texture-coordinate derivatives (Lengyel 2001), float32:
    per triangle dP/du and dP/dv summed at its three vertices in triangle order, made orthogonal to the vertex normal
    (the given one, or the sum of the faces' cross(e1, e2)); handedness +1 / -1 from the sign of cross(n, t) . dP/dv;
    a vertex with no usable derivative gets cross(a, n) normalised, a = x unless |n.x| >= 0.9, then y."""
    v = np.asarray(verts, np.float32)
    f = np.asarray(tris, np.int64)
    q = np.asarray(uv, np.float32).reshape(-1, 6)
    nv = len(v)

    def dot(a, b):
        return a[:, 0] * b[:, 0] + a[:, 1] * b[:, 1] + a[:, 2] * b[:, 2]

    def cross(a, b):
        return np.stack([a[:, 1] * b[:, 2] - a[:, 2] * b[:, 1], a[:, 2] * b[:, 0] - a[:, 0] * b[:, 2],
                         a[:, 0] * b[:, 1] - a[:, 1] * b[:, 0]], 1)

    def normalize(a):
        ln = np.sqrt(dot(a, a))
        inv = np.where(ln > 0, np.float32(1) / np.where(ln > 0, ln, np.float32(1)), np.float32(1)).astype(np.float32)
        return (a * inv[:, None]).astype(np.float32)

    e1, e2 = v[f[:, 1]] - v[f[:, 0]], v[f[:, 2]] - v[f[:, 0]]
    du1, dv1, du2, dv2 = q[:, 2] - q[:, 0], q[:, 3] - q[:, 1], q[:, 4] - q[:, 0], q[:, 5] - q[:, 1]
    det = du1 * dv2 - du2 * dv1
    ok = np.abs(det) >= np.float32(1e-20)
    r = (np.float32(1) / np.where(ok, det, np.float32(1))).astype(np.float32)
    dpdu = ((e1 * dv2[:, None] - e2 * dv1[:, None]) * r[:, None]).astype(np.float32)
    dpdv = ((e2 * du1[:, None] - e1 * du2[:, None]) * r[:, None]).astype(np.float32)
    # sums in the order triangle 0 corners 0, 1, 2, triangle 1 corners 0, 1, 2, ...
    su, sv = np.zeros((nv, 3), np.float32), np.zeros((nv, 3), np.float32)
    fo = f[ok].reshape(-1)
    np.add.at(su, fo, np.repeat(dpdu[ok], 3, axis=0))
    np.add.at(sv, fo, np.repeat(dpdv[ok], 3, axis=0))
    if normals is not None and len(normals) == nv:
        n = np.asarray(normals, np.float32)
    else:
        n = np.zeros((nv, 3), np.float32)
        np.add.at(n, f.reshape(-1), np.repeat(cross(e1, e2).astype(np.float32), 3, axis=0))
    n = normalize(n)
    tg = (su - n * dot(n, su)[:, None]).astype(np.float32)
    bad = dot(tg, tg) < np.float32(1e-24)
    a = np.where((np.abs(n[:, 0]) < np.float32(0.9))[:, None], np.float32([1, 0, 0]), np.float32([0, 1, 0])).astype(np.float32)
    tg[bad] = cross(a[bad], n[bad])
    tg = normalize(tg)
    w = np.where(dot(cross(n, tg), sv) < 0, np.float32(-1), np.float32(1)).astype(np.float32)
    per = np.concatenate([tg, w[:, None]], 1)
    return per[f].astype(np.float32)


def texture_sample(img, texels, u, v, nearest=False, table=None):
    """The oracle's texture lookup (rgba)."""
    im = np.ascontiguousarray(img, np.int32)
    tx = np.ascontiguousarray(texels, np.uint32)
    tab = srgb_table() if table is None else np.ascontiguousarray(table, np.float32)
    out = np.zeros(4, np.float32)
    lib().orc_texture_sample(_i(im), tx.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)), _f(tab), float(u), float(v),
                             1 if nearest else 0, _f(out))
    return out


def display(acc, spp, exposure_ev=0.0, mode=0):
    """the viewer's display transform of (n, 4) float32 sums -> (n, 4) uint8 RGBA (src/app/display.cu:9-33)"""
    a = np.ascontiguousarray(acc, np.float32).reshape(-1, 4)
    out = np.zeros((a.shape[0], 4), np.uint8)
    lib().orc_display(_f(a), a.shape[0], int(spp), float(exposure_ev), int(mode),
                      out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)))
    return out


def srgb_encode8(x):
    return lib().orc_srgb_encode8(float(x))


def srgb_steps_check(steps, lo=0.0, hi=1.0, threads=8):
    """number of floats in [lo, hi] whose 8-bit encode differs from the search over the 255 step positions"""
    t = np.ascontiguousarray(steps, np.float32)
    assert t.shape == (256,)
    return int(lib().orc_srgb_steps_check(_f(t), float(lo), float(hi), int(threads)))
