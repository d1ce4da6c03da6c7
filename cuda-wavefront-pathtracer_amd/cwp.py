"""Python binding of libcwp.so (the C ABI in include/cwp.h) over ctypes.

The library is the product: this module only marshals arguments. It raises if libcwp.so is missing; there is no
fallback implementation.

    import cwp
    s = cwp.Setup(["--gen", "terrain", "--width", "1280", "--height", "720", "--batch", "256"])
    r = cwp.Renderer(s)
    stats = r.render(256)
    img = r.accum() / 256
"""
import ctypes
import json
import os

import numpy as np

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "libcwp.so")

ARR_VERTICES, ARR_TRIANGLES, ARR_MATERIAL, ARR_ALBEDO, ARR_NODES, ARR_ORDER, ARR_WOOP, ARR_CORNERS, ARR_NORMALS, \
    ARR_TREE, ARR_TREE_BOX, ARR_PRIM, ARR_MESH_ROOTS, ARR_MESH_BOXES, ARR_MATERIALS, ARR_EM_TRI, ARR_EM_CDF, \
    ARR_TRI_LIGHT_PDF, ARR_ENV_TEX, ARR_ENV_MARG, ARR_ENV_COND, ARR_MESH_TRIS, ARR_UVS, ARR_TANGENTS, ARR_MAPS, \
    ARR_TEX_IMAGES, ARR_TEXELS, ARR_SRGB_TABLE, ARR_MESH_UVS = range(29)
TRI_MT, TRI_WOOP, TRI_WATERTIGHT = 0, 1, 2
CWP_COPY_H2D, CWP_COPY_D2H, CWP_COPY_D2D = 1, 2, 3


class CwpError(RuntimeError):
    pass


class Params(ctypes.Structure):
    """cwp_params: frame parameters of the legacy integrator."""
    _fields_ = [("pos", ctypes.c_float * 3), ("ax", ctypes.c_float * 3), ("ay", ctypes.c_float * 3),
                ("az", ctypes.c_float * 3), ("fx", ctypes.c_float), ("fy", ctypes.c_float), ("px", ctypes.c_float),
                ("py", ctypes.c_float), ("w", ctypes.c_int), ("h", ctypes.c_int), ("lens_r", ctypes.c_float),
                ("focus_d", ctypes.c_float), ("sun_to", ctypes.c_float * 3), ("sun_e", ctypes.c_float * 3),
                ("sky", ctypes.c_float * 3), ("sun_cos", ctypes.c_float), ("max_bounce", ctypes.c_int),
                ("seed", ctypes.c_uint32), ("eps", ctypes.c_float), ("near_t", ctypes.c_float),
                ("has_normals", ctypes.c_int), ("filter_bh", ctypes.c_int), ("ftab", ctypes.c_float * 65)]


class FrameStats(ctypes.Structure):
    _fields_ = [("rays_closest", ctypes.c_double), ("rays_shadow", ctypes.c_double), ("t_trace_ms", ctypes.c_double),
                ("t_frame_ms", ctypes.c_double), ("t_bounce_ms", ctypes.c_double), ("t_shadow_ms", ctypes.c_double),
                ("bounce_launches", ctypes.c_int), ("shadow_launches", ctypes.c_int), ("rays_hit", ctypes.c_double),
                ("rays_miss", ctypes.c_double), ("rays_cont", ctypes.c_double), ("shadow_unoccluded", ctypes.c_double),
                ("camera_miss", ctypes.c_double), ("t_shade_ms", ctypes.c_double), ("t_sort_ms", ctypes.c_double),
                ("t_tail_ms", ctypes.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class SceneInfo(ctypes.Structure):
    _fields_ = [("variant", ctypes.c_int), ("tri_test", ctypes.c_int), ("num_tris", ctypes.c_int),
                ("num_verts", ctypes.c_int), ("num_materials", ctypes.c_int), ("num_nodes", ctypes.c_int),
                ("node_floats4", ctypes.c_int), ("bvh_depth", ctypes.c_int), ("stack_need", ctypes.c_int),
                ("sah", ctypes.c_double), ("bvh_build_s", ctypes.c_double), ("width", ctypes.c_int),
                ("height", ctypes.c_int), ("spp", ctypes.c_int), ("batch", ctypes.c_int), ("bounces", ctypes.c_int),
                ("num_meshes", ctypes.c_int), ("num_instances", ctypes.c_int), ("frames", ctypes.c_int),
                ("frame_start", ctypes.c_int), ("full", ctypes.c_int), ("num_emitters", ctypes.c_int),
                ("env_w", ctypes.c_int), ("env_h", ctypes.c_int), ("sun_on", ctypes.c_int), ("light_mode", ctypes.c_int),
                ("rr", ctypes.c_int), ("rr_depth", ctypes.c_int), ("gd_energy", ctypes.c_int),
                ("env_uniform", ctypes.c_int), ("env_rot_deg", ctypes.c_float),
                ("bvh_cache", ctypes.c_int), ("textured", ctypes.c_int), ("num_images", ctypes.c_int),
                ("tex_nearest", ctypes.c_int), ("bump_shadow", ctypes.c_int)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class Instance(ctypes.Structure):
    """cwp_instance: world-to-object rows, object-to-world rows, BLAS root, mesh, det sign."""
    _fields_ = [("inv", ctypes.c_float * 12), ("fwd", ctypes.c_float * 12), ("root", ctypes.c_int), ("mesh", ctypes.c_int),
                ("det_sign", ctypes.c_float), ("pad", ctypes.c_int)]


class SceneDesc(ctypes.Structure):
    """cwp_scene_desc: a scene in the renderer's upload layout (host arrays)."""
    _fields_ = [("variant", ctypes.c_int), ("nodes", ctypes.c_void_p), ("num_nodes", ctypes.c_size_t),
                ("corners", ctypes.c_void_p), ("num_tris", ctypes.c_size_t), ("woop", ctypes.c_void_p),
                ("normals", ctypes.c_void_p), ("albedo", ctypes.c_void_p), ("num_materials", ctypes.c_size_t),
                ("tri_form", ctypes.c_int), ("root", ctypes.c_int),
                ("mats", ctypes.c_void_p), ("tri_light_pdf", ctypes.c_void_p), ("em_tri", ctypes.c_void_p),
                ("em_cdf", ctypes.c_void_p), ("num_emitters", ctypes.c_size_t), ("env_texels", ctypes.c_void_p),
                ("env_w", ctypes.c_int), ("env_h", ctypes.c_int), ("env_rot_deg", ctypes.c_float), ("sun_on", ctypes.c_int),
                ("light_mode", ctypes.c_int), ("rr", ctypes.c_int), ("rr_depth", ctypes.c_int), ("gd_energy", ctypes.c_int),
                ("env_uniform", ctypes.c_int), ("maps", ctypes.c_void_p), ("uvs", ctypes.c_void_p),
                ("tangents", ctypes.c_void_p), ("tex_images", ctypes.c_void_p), ("num_images", ctypes.c_size_t),
                ("texels", ctypes.c_void_p), ("num_texels", ctypes.c_size_t), ("tex_nearest", ctypes.c_int),
                ("bump_shadow", ctypes.c_int)]


class Options(ctypes.Structure):
    """cwp_options: kernel and schedule options."""
    _fields_ = [("tri_test", ctypes.c_int), ("batch", ctypes.c_int), ("bounces", ctypes.c_int), ("fuse", ctypes.c_int),
                ("persistent", ctypes.c_int), ("shadow_refill", ctypes.c_int), ("stack_lds", ctypes.c_int),
                ("blocks_per_cu", ctypes.c_int), ("bounce_waves", ctypes.c_int), ("time_kernels", ctypes.c_int),
                ("graph", ctypes.c_int), ("sort", ctypes.c_int), ("tail_lanes", ctypes.c_int)]

    @classmethod
    def default(cls, **kw):
        o = cls()
        lib().cwp_options_default(ctypes.byref(o))
        for k, v in kw.items():
            setattr(o, k, v)
        return o


class ServerOptions(ctypes.Structure):
    """cwp_server_options: the server's BVH layout, triangle test, LDS stack, bounce limit, path-state slots, jobs per
    pass, scene budget and mode."""
    _fields_ = [("variant", ctypes.c_int), ("tri_test", ctypes.c_int), ("stack_lds", ctypes.c_int), ("bounces", ctypes.c_int),
                ("max_paths", ctypes.c_size_t), ("max_jobs", ctypes.c_int), ("scene_budget_mb", ctypes.c_double),
                ("batched", ctypes.c_int), ("rays_per_lane", ctypes.c_int), ("use_graph", ctypes.c_int),
                ("full", ctypes.c_int), ("rolling", ctypes.c_int), ("compact", ctypes.c_int), ("tail_lanes", ctypes.c_int),
                ("tail_levels", ctypes.c_int)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}

    @classmethod
    def default(cls, **kw):
        o = cls()
        lib().cwp_server_options_default(ctypes.byref(o))
        for k, v in kw.items():
            setattr(o, k, v)
        return o


class TuneTrial(ctypes.Structure):
    _fields_ = [("variant", ctypes.c_int), ("stack_lds", ctypes.c_int), ("rays_per_lane", ctypes.c_int),
                ("compact", ctypes.c_int), ("rays_per_s", ctypes.c_double), ("wall_s", ctypes.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class JobInfo(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int) for n in ("scene", "width", "height", "spp", "bounces", "samples_per_pass", "passes")] + \
               [(n, ctypes.c_double) for n in ("arrival_s", "admitted_s", "finished_s", "latency_s")]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class ServeStats(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int) for n in ("jobs", "passes", "scenes_loaded", "scenes_evicted")] + \
               [(n, ctypes.c_double) for n in ("wall_s", "rays", "rays_per_s", "latency_p50_s", "latency_p99_s",
                                               "latency_mean_s", "max_paths_in_flight")] + [("steps", ctypes.c_longlong)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


def scene_desc(variant, nodes, corners, albedo, woop=None, normals=None, tri_form=0, root=0, full=None):
    """Builds a cwp_scene_desc over numpy arrays; returns (desc, keep-alive list). tri_form 1: corners holds the
    reference's edge form (v0, e1, e2). full: a dict of the full integrator's fields (mats (m, 16) float32,
    tri_light_pdf, em_tri, em_cdf, env_texels (h, w, 4), env_rot_deg, sun_on, light_mode, rr, rr_depth, gd_energy,
    env_uniform; textures: maps (m, 8) int32, uvs (n, 2, 4), tangents (n, 3, 4), tex_images (k, 4) int32, texels uint32,
    tex_nearest, bump_shadow), or None for a legacy-integrator scene."""
    keep = [np.ascontiguousarray(a, dtype=np.float32) if a is not None else None
            for a in (nodes, corners, woop, normals, albedo)]
    n, c, w, nr, al = keep
    per = {2: 16, 4: 32, 8: 20}[variant]
    d = SceneDesc(variant, n.ctypes.data, n.size // per, c.ctypes.data, c.size // 12,
                  w.ctypes.data if w is not None else None, nr.ctypes.data if nr is not None else None,
                  al.ctypes.data, al.size // 4, tri_form, root)
    if full is not None:
        def arr(key, dt):
            a = full.get(key)
            if a is None or len(a) == 0:
                return None
            a = np.ascontiguousarray(a, dtype=dt)
            keep.append(a)
            return a
        m = arr("mats", np.float32)
        d.mats = m.ctypes.data
        tlp, emt, emc = arr("tri_light_pdf", np.float32), arr("em_tri", np.int32), arr("em_cdf", np.float32)
        if tlp is not None:
            d.tri_light_pdf, d.em_tri, d.em_cdf, d.num_emitters = tlp.ctypes.data, emt.ctypes.data, emc.ctypes.data, len(emt)
        env = arr("env_texels", np.float32)
        if env is not None:
            d.env_texels, d.env_h, d.env_w = env.ctypes.data, env.shape[0], env.shape[1]
        d.env_rot_deg = full.get("env_rot_deg", 0.0)
        d.sun_on = full.get("sun_on", 1)
        d.light_mode = full.get("light_mode", 0)
        d.rr, d.rr_depth = full.get("rr", 0), full.get("rr_depth", 3)
        d.gd_energy, d.env_uniform = full.get("gd_energy", 1), full.get("env_uniform", 0)
        maps = arr("maps", np.int32)
        if maps is not None:
            uv, tg, ti, tx = arr("uvs", np.float32), arr("tangents", np.float32), arr("tex_images", np.int32), arr("texels", np.uint32)
            d.maps, d.uvs = maps.ctypes.data, uv.ctypes.data
            d.tangents = tg.ctypes.data if tg is not None else None
            d.tex_images, d.num_images = ti.ctypes.data, len(ti)
            d.texels, d.num_texels = tx.ctypes.data, tx.size
            d.tex_nearest, d.bump_shadow = full.get("tex_nearest", 0), full.get("bump_shadow", 0)
    return d, keep


def scene_check(variant, nodes, corners, albedo, woop=None, normals=None, tri_form=0, root=0):
    """Host-side check of a scene description; returns the traversal-stack entries the tree can need."""
    d, keep = scene_desc(variant, nodes, corners, albedo, woop, normals, tri_form, root)
    need = ctypes.c_int(0)
    _check(lib().cwp_scene_check(ctypes.byref(d), ctypes.byref(need)))
    return need.value


# declared symbols of include/cwp.h with their signatures
_P, _I, _D, _S = ctypes.c_void_p, ctypes.c_int, ctypes.c_double, ctypes.c_size_t
_F, _FP = ctypes.c_float, ctypes.POINTER(ctypes.c_float)
SIGNATURES = {
    "cwp_last_error": ([], ctypes.c_char_p),
    "cwp_version": ([], _I),
    "cwp_device_info": ([ctypes.POINTER(_I), ctypes.c_char_p, _S], _I),
    "cwp_setup_create": ([_I, ctypes.POINTER(ctypes.c_char_p), ctypes.POINTER(_P)], _I),
    "cwp_setup_destroy": ([_P], None),
    "cwp_setup_info": ([_P, ctypes.POINTER(SceneInfo)], _I),
    "cwp_setup_params": ([_P, ctypes.POINTER(Params)], _I),
    "cwp_setup_array": ([_P, _I, ctypes.POINTER(_P), ctypes.POINTER(_S)], _I),
    "cwp_create": ([_P, ctypes.POINTER(_P)], _I),
    "cwp_destroy": ([_P], None),
    "cwp_render": ([_P, _I, _I, _I, ctypes.POINTER(FrameStats)], _I),
    "cwp_read_accum": ([_P, ctypes.POINTER(ctypes.c_float), _S], _I),
    "cwp_set_params": ([_P, ctypes.POINTER(Params)], _I),
    "cwp_accum_device": ([_P], _P),
    "cwp_display": ([_P, _P, _I, ctypes.c_float, _I], _I),
    "cwp_display_buffer": ([_P, _P, _I, _I, ctypes.c_float, _I], _I),
    "cwp_srgb_encode8": ([ctypes.c_float], _I),
    "cwp_srgb_steps": ([ctypes.POINTER(ctypes.c_float)], None),
    "cwp_set_materials": ([_P, ctypes.POINTER(ctypes.c_float), _I], _I),
    "cwp_set_environment": ([_P, ctypes.POINTER(ctypes.c_float), _S, ctypes.c_float, _I], _I),
    "cwp_trace": ([_P, _P, _I, _I, _P], _I),
    "cwp_trace_host": ([_P, ctypes.POINTER(ctypes.c_float), _I, _I, ctypes.POINTER(ctypes.c_float)], _I),
    "cwp_frame_json": ([_P, _P, ctypes.POINTER(FrameStats), _I, _I, ctypes.c_char_p, _S], _I),
    "cwp_ctx_info": ([_P, ctypes.POINTER(_I), ctypes.POINTER(_D)], _I),
    "cwp_options_default": ([ctypes.POINTER(Options)], None),
    "cwp_setup_instances": ([_P, _I, ctypes.POINTER(Instance), _I, ctypes.POINTER(_I)], _I),
    "cwp_set_instances": ([_P, ctypes.POINTER(Instance), _I, ctypes.POINTER(_D)], _I),
    "cwp_set_frame": ([_P, _I, ctypes.POINTER(_D)], _I),
    "cwp_trace_host_inst": ([_P, ctypes.POINTER(ctypes.c_float), _I, _I, ctypes.POINTER(ctypes.c_float),
                             ctypes.POINTER(_I)], _I),
    "cwp_scene_check": ([ctypes.POINTER(SceneDesc), ctypes.POINTER(_I)], _I),
    "cwp_create_scene": ([ctypes.POINTER(SceneDesc), ctypes.POINTER(Params), ctypes.POINTER(Options), ctypes.POINTER(_P)], _I),
    "cwp_bvh_build": ([ctypes.POINTER(ctypes.c_float), _I, ctypes.POINTER(_I), _I, _I, _I, _I, ctypes.POINTER(_P)], _I),
    "cwp_bvh_destroy": ([_P], None),
    "cwp_bvh_array": ([_P, _I, ctypes.POINTER(_P), ctypes.POINTER(_S)], _I),
    "cwp_bvh_stats": ([_P, ctypes.POINTER(_I), ctypes.POINTER(_I), ctypes.POINTER(_D)], _I),
    "cwp_bvh8_device_nodes": ([ctypes.POINTER(ctypes.c_float), _S, ctypes.POINTER(ctypes.c_float)], _I),
    "cwp_bvh_from_hploc": ([ctypes.POINTER(ctypes.c_float), _I, ctypes.POINTER(_I), _I, ctypes.POINTER(ctypes.c_float), _I,
                            _I, _I, ctypes.POINTER(_P)], _I),
    "cwp_hploc_build_host": ([ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_float), _I,
                              ctypes.POINTER(ctypes.c_float), ctypes.POINTER(_I), ctypes.POINTER(_D)], _I),
    "cwp_hploc_build_host_mode": ([_FP, _FP, _I, _I, _FP, ctypes.POINTER(_I), ctypes.POINTER(_D)], _I),
    "cwp_hploc_build_mode": ([_P, _P, _P, _I, _I, _P, ctypes.POINTER(_I)], _I),
    "cwp_hploc_build_static": ([_P, _P, _I, _I, _P, _P, ctypes.POINTER(_I)], _I),
    "cwp_hploc_create": ([_I, ctypes.POINTER(_P)], _I),
    "cwp_hploc_destroy": ([_P], None),
    "cwp_hploc_build": ([_P, _P, _P, _I, _P, ctypes.POINTER(_I)], _I),
    "cwp_hploc_nodes": ([_P, ctypes.POINTER(_P)], _I),
    "cwp_hploc_to_bvh2": ([_P, _I, _P, _P], _I),
    "cwp_cwbvh_from_hploc": ([_P, _I, _P, ctypes.POINTER(_I), _P, ctypes.POINTER(_I), _P], _I),
    "cwp_cwbvh_from_hploc_host": ([ctypes.POINTER(ctypes.c_float), _I, ctypes.POINTER(ctypes.c_float), ctypes.POINTER(_I),
                                   ctypes.POINTER(_I), ctypes.POINTER(_I), ctypes.POINTER(_D)], _I),
    "cwp_set_emitters": ([_P, ctypes.POINTER(ctypes.c_float), _I, ctypes.POINTER(ctypes.c_float),
                          ctypes.POINTER(ctypes.c_float), _I], _I),
    "cwp_setup_emitters": ([_P, _I, ctypes.POINTER(ctypes.c_float), _I, ctypes.POINTER(_I), ctypes.POINTER(ctypes.c_float),
                            ctypes.POINTER(ctypes.c_float), _I], _I),
    "cwp_server_options_default": ([ctypes.POINTER(ServerOptions)], None),
    "cwp_server_create": ([ctypes.POINTER(ServerOptions), ctypes.POINTER(_P)], _I),
    "cwp_server_destroy": ([_P], None),
    "cwp_server_add_scene": ([_P, _P, ctypes.POINTER(_I), ctypes.POINTER(_S)], _I),
    "cwp_server_submit": ([_P, _I, _I, _I, _I, _I, ctypes.c_uint32, _D, ctypes.POINTER(_I)], _I),
    "cwp_server_run": ([_P, _I], _I),
    "cwp_server_job": ([_P, _I, ctypes.POINTER(JobInfo)], _I),
    "cwp_server_image": ([_P, _I, ctypes.POINTER(ctypes.c_float), _S], _I),
    "cwp_server_stats": ([_P, ctypes.POINTER(ServeStats)], _I),
    "cwp_server_clear_jobs": ([_P], _I),
    "cwp_server_reset_jobs": ([_P], _I),
    "cwp_server_autotune": ([_P, ctypes.POINTER(_I), _I, ctypes.POINTER(_I), _I, ctypes.POINTER(_I), _I,
                             ctypes.POINTER(_I), _I, ctypes.POINTER(TuneTrial), _I, ctypes.POINTER(_I),
                             ctypes.POINTER(ServerOptions)], _I),
    "cwp_server_get_options": ([_P, ctypes.POINTER(ServerOptions)], _I),
    "cwp_server_preload": ([_P], _I),
    "cwp_device_malloc": ([_S, ctypes.POINTER(_P)], _I),
    "cwp_device_free": ([_P], None),
    "cwp_memcpy": ([_P, _P, _S, _I], _I),
    "cwp_bsdf_eval": ([_FP, _FP, _FP, _F, _I, _F, _F, _FP, _FP], _I),
    "cwp_image_read": ([ctypes.c_char_p, ctypes.POINTER(_I), ctypes.POINTER(_I), _P, _S], _I),
    "cwp_inflate": ([_P, _S, _P, _S, ctypes.POINTER(_S)], _I),
    "cwp_texture_sample": ([ctypes.POINTER(_I), _P, _S, _F, _F, _I, _FP], _I),
    "cwp_server_start": ([_P], _I),
    "cwp_job_submit": ([_P, _I, _I, _I, _I, _I, ctypes.c_uint32, ctypes.POINTER(_I)], _I),
    "cwp_job_poll": ([_P, _I, ctypes.POINTER(_I)], _I),
    "cwp_job_result": ([_P, _I, ctypes.POINTER(ctypes.c_float), _S], _I),
    "cwp_server_stop": ([_P], _I),
    "cwp_bsdf_sample": ([_FP, _FP, _F, _I, _F, _F, _F, _F, _F, _FP, _FP, _FP, ctypes.POINTER(_I)], _I),
}

_lib = None


def lib():
    """Loads libcwp.so once (raises CwpError if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise CwpError("libcwp.so not found at %s: build it with cuda-wavefront-pathtracer_amd/build.py" % LIB_PATH)
        L = ctypes.CDLL(LIB_PATH)
        for name, (args, res) in SIGNATURES.items():
            try:
                f = getattr(L, name)
            except AttributeError:   # an older build: calling this entry point fails, the others work
                continue
            f.argtypes = args
            f.restype = res
        _lib = L
    return _lib


def _check(rc):
    if rc != 0:
        raise CwpError("cwp error %d: %s" % (rc, lib().cwp_last_error().decode()))


def device_info():
    n = ctypes.c_int(0)
    name = ctypes.create_string_buffer(256)
    _check(lib().cwp_device_info(ctypes.byref(n), name, 256))
    return n.value, name.value.decode()


# element dtype and shape per element of each array selector
_ARRAY_LAYOUT = {ARR_VERTICES: (np.float32, (3,)), ARR_TRIANGLES: (np.int32, (3,)), ARR_MATERIAL: (np.int32, ()),
                 ARR_ALBEDO: (np.float32, (4,)), ARR_ORDER: (np.int32, ()), ARR_WOOP: (np.float32, (3, 4)),
                 ARR_CORNERS: (np.float32, (3, 4)), ARR_NORMALS: (np.float32, (3,)), ARR_TREE: (np.int32, (4,)),
                 ARR_TREE_BOX: (np.float32, (6,)), ARR_PRIM: (np.int32, ()), ARR_MESH_ROOTS: (np.int32, ()),
                 ARR_MESH_BOXES: (np.float32, (8,)), ARR_MATERIALS: (np.float32, (16,)), ARR_EM_TRI: (np.int32, ()),
                 ARR_EM_CDF: (np.float32, ()), ARR_TRI_LIGHT_PDF: (np.float32, ()), ARR_ENV_TEX: (np.float32, (4,)),
                 ARR_ENV_MARG: (np.float32, ()), ARR_ENV_COND: (np.float32, ()), ARR_MESH_TRIS: (np.int32, (2,)),
                 ARR_UVS: (np.float32, (2, 4)), ARR_TANGENTS: (np.float32, (3, 4)), ARR_MAPS: (np.int32, (8,)),
                 ARR_TEX_IMAGES: (np.int32, (4,)), ARR_TEXELS: (np.uint32, ()), ARR_SRGB_TABLE: (np.float32, ()),
                 ARR_MESH_UVS: (np.float32, (6,))}


def _view(ptr, count, which, node_floats4=None):
    if which == ARR_NODES:
        dtype, shape = np.float32, (node_floats4 * 4,)
    else:
        dtype, shape = _ARRAY_LAYOUT[which]
    n = count * int(np.prod(shape)) if shape else count
    if n == 0 or not ptr:
        return np.zeros((0,) + shape, dtype=dtype)
    buf = (ctypes.c_char * (n * np.dtype(dtype).itemsize)).from_address(ptr)
    return np.frombuffer(buf, dtype=dtype).reshape((count,) + shape).copy()


class Setup:
    """Parsed options + scene + BVH (host). args: the pt command-line flags (without argv[0])."""

    def __init__(self, args):
        argv = [b"cwp"] + [str(a).encode() for a in args]
        arr = (ctypes.c_char_p * len(argv))(*argv)
        h = ctypes.c_void_p()
        _check(lib().cwp_setup_create(len(argv), arr, ctypes.byref(h)))
        self.h = h
        self.info = self._info()

    def _info(self):
        i = SceneInfo()
        _check(lib().cwp_setup_info(self.h, ctypes.byref(i)))
        return i.as_dict()

    def params(self):
        p = Params()
        _check(lib().cwp_setup_params(self.h, ctypes.byref(p)))
        return p

    def array(self, which):
        ptr, cnt = ctypes.c_void_p(), ctypes.c_size_t()
        _check(lib().cwp_setup_array(self.h, which, ctypes.byref(ptr), ctypes.byref(cnt)))
        return _view(ptr.value, cnt.value, which, self.info["node_floats4"])

    def instances(self, frame):
        """Instance records of a frame of a two-level scene: (n, 28) float32 (inv rows, fwd rows, root / mesh bits,
        det sign, pad)."""
        n = ctypes.c_int()
        _check(lib().cwp_setup_instances(self.h, frame, None, 0, ctypes.byref(n)))
        rec = (Instance * n.value)()
        _check(lib().cwp_setup_instances(self.h, frame, rec, n.value, ctypes.byref(n)))
        return np.frombuffer(rec, dtype=np.float32).reshape(n.value, 28).copy()

    def emitters(self, frame):
        """Two-level full-integrator scenes: (world emitters (n, 3, 4), power CDF (n + 1), per-material light pdf)."""
        n = ctypes.c_int()
        _check(lib().cwp_setup_emitters(self.h, frame, None, 0, ctypes.byref(n), None, None, 0))
        em = np.zeros((n.value, 3, 4), np.float32)
        cdf = np.zeros(n.value + 1, np.float32)
        mlp = np.zeros(self.info["num_materials"], np.float32)
        fp = ctypes.POINTER(ctypes.c_float)
        _check(lib().cwp_setup_emitters(self.h, frame, em.ctypes.data_as(fp), n.value, ctypes.byref(n), cdf.ctypes.data_as(fp),
                                        mlp.ctypes.data_as(fp), len(mlp)))
        return em, cdf, mlp

    def close(self):
        if getattr(self, "h", None):
            lib().cwp_setup_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Renderer:
    """Device renderer of a Setup (uploads the scene, allocates path state)."""

    def __init__(self, setup):
        self.setup = setup
        h = ctypes.c_void_p()
        _check(lib().cwp_create(setup.h, ctypes.byref(h)))
        self._attach(h, setup.info["width"], setup.info["height"])

    def _attach(self, h, w, hgt):
        self.h = h
        self.w, self.hgt = w, hgt
        g, v = ctypes.c_int(), ctypes.c_double()
        _check(lib().cwp_ctx_info(self.h, ctypes.byref(g), ctypes.byref(v)))
        self.grid_blocks, self.vram_mb = g.value, v.value

    @classmethod
    def from_scene(cls, variant, nodes, corners, albedo, params, options, woop=None, normals=None, tri_form=0, root=0,
                   full=None):
        """Renderer over caller arrays in the upload layout (cwp_create_scene); full: see scene_desc."""
        d, keep = scene_desc(variant, nodes, corners, albedo, woop, normals, tri_form, root, full)
        h = ctypes.c_void_p()
        _check(lib().cwp_create_scene(ctypes.byref(d), ctypes.byref(params), ctypes.byref(options), ctypes.byref(h)))
        r = cls.__new__(cls)
        r.setup = None
        r._attach(h, params.w, params.h)
        return r

    def render(self, spp, sample_base=0, clear=True):
        st = FrameStats()
        _check(lib().cwp_render(self.h, int(spp), int(sample_base), 1 if clear else 0, ctypes.byref(st)))
        self._last = st
        return st.as_dict()

    def accum(self):
        w, h = self.w, self.hgt
        out = np.zeros((h, w, 4), dtype=np.float32)
        _check(lib().cwp_read_accum(self.h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), w * h))
        return out

    def accum_device(self):
        """device address of the accumulation buffer (w * h float4 sums)"""
        return lib().cwp_accum_device(self.h)

    def display(self, spp, exposure_ev=0.0, mode="clip", out=None):
        """The viewer's display transform of the accumulation on the GPU: (h, w, 4) uint8 RGBA. out: an optional
        DeviceArray of w * h * 4 bytes that receives the pixels (and is then also read back)."""
        n = self.w * self.hgt
        buf = out if out is not None else DeviceArray(n * 4)
        if buf.nbytes < n * 4:
            raise ValueError("display buffer too small")
        _check(lib().cwp_display(self.h, buf.ptr, int(spp), float(exposure_ev), DISPLAY_MODES[mode]))
        return buf.to_host(np.uint8, (self.hgt, self.w, 4))

    def set_params(self, params):
        _check(lib().cwp_set_params(self.h, ctypes.byref(params)))

    def set_materials(self, mats):
        """mats: (m, 16) float32 material records (the scene's count)."""
        m = np.ascontiguousarray(mats, np.float32).reshape(-1, 16)
        _check(lib().cwp_set_materials(self.h, m.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), len(m)))

    def set_environment(self, texels=None, rot_deg=0.0, sun_on=True):
        """texels: (h, w, 4) float32 radiance of the scene's size, or None to keep them."""
        if texels is None:
            _check(lib().cwp_set_environment(self.h, None, 0, float(rot_deg), 1 if sun_on else 0))
            return
        t = np.ascontiguousarray(texels, np.float32).reshape(-1, 4)
        _check(lib().cwp_set_environment(self.h, t.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), len(t), float(rot_deg),
                                         1 if sun_on else 0))

    def trace_ptr(self, rays_ptr, n, any_hit, hits_ptr):
        """rays_ptr / hits_ptr: device addresses (e.g. torch.Tensor.data_ptr())."""
        _check(lib().cwp_trace(self.h, ctypes.c_void_p(rays_ptr), int(n), 1 if any_hit else 0, ctypes.c_void_p(hits_ptr)))

    def set_frame(self, frame):
        """Two-level scenes: instances of a frame of the setup, TLAS rebuilt on the GPU; returns the TLAS build ms."""
        ms = ctypes.c_double()
        _check(lib().cwp_set_frame(self.h, int(frame), ctypes.byref(ms)))
        return ms.value

    def set_instances(self, records):
        """records: (n, 28) float32 instance records; returns the TLAS build ms."""
        r = np.ascontiguousarray(records, dtype=np.float32).reshape(-1, 28)
        ms = ctypes.c_double()
        _check(lib().cwp_set_instances(self.h, ctypes.cast(r.ctypes.data, ctypes.POINTER(Instance)), len(r), ctypes.byref(ms)))
        return ms.value

    def set_emitters(self, em_tris, em_cdf, mat_light_pdf):
        fp = ctypes.POINTER(ctypes.c_float)
        e = np.ascontiguousarray(em_tris, np.float32).reshape(-1, 3, 4)
        c = np.ascontiguousarray(em_cdf, np.float32)
        m = np.ascontiguousarray(mat_light_pdf, np.float32)
        _check(lib().cwp_set_emitters(self.h, e.ctypes.data_as(fp), len(e), c.ctypes.data_as(fp), m.ctypes.data_as(fp), len(m)))

    def trace_inst(self, rays, any_hit=False):
        """Two-level scenes: (hits (n, 4), instance ids (n,))."""
        r = np.ascontiguousarray(rays, dtype=np.float32).reshape(-1, 8)
        hits = np.zeros((r.shape[0], 4), dtype=np.float32)
        ids = np.zeros(r.shape[0], dtype=np.int32)
        fp = ctypes.POINTER(ctypes.c_float)
        _check(lib().cwp_trace_host_inst(self.h, r.ctypes.data_as(fp), r.shape[0], 1 if any_hit else 0,
                                         hits.ctypes.data_as(fp), ids.ctypes.data_as(ctypes.POINTER(ctypes.c_int))))
        return hits, ids

    def trace(self, rays, any_hit=False):
        """rays: (n, 8) float32 array (o.xyz, tmin, d.xyz, tmax). Returns (n, 4) hits (t, u, v, leaf position bits)."""
        r = np.ascontiguousarray(rays, dtype=np.float32).reshape(-1, 8)
        hits = np.zeros((r.shape[0], 4), dtype=np.float32)
        fp = ctypes.POINTER(ctypes.c_float)
        _check(lib().cwp_trace_host(self.h, r.ctypes.data_as(fp), r.shape[0], 1 if any_hit else 0, hits.ctypes.data_as(fp)))
        return hits

    def frame_json(self, stats=None, rep=0, frame=0):
        st = self._last if stats is None else stats
        if isinstance(st, dict):
            f = FrameStats()
            for k, v in st.items():
                setattr(f, k, v)
            st = f
        buf = ctypes.create_string_buffer(8192)
        sh = self.setup.h if self.setup is not None else None
        _check(lib().cwp_frame_json(sh, self.h, ctypes.byref(st), rep, frame, buf, 8192))
        return json.loads(buf.value.decode())

    def close(self):
        if getattr(self, "h", None):
            lib().cwp_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Server:
    """Render-job server (cwp_server_*): scenes from Setups, jobs with arrival times, batched or plain runs."""

    def __init__(self, options=None, **kw):
        o = options if options is not None else ServerOptions.default(**kw)
        h = ctypes.c_void_p()
        _check(lib().cwp_server_create(ctypes.byref(o), ctypes.byref(h)))
        self.h = h
        self.options = o

    def add_scene(self, setup):
        """Returns the scene id; scene_bytes[id] is the device bytes it takes when resident."""
        sid, nb = ctypes.c_int(), ctypes.c_size_t()
        _check(lib().cwp_server_add_scene(self.h, setup.h, ctypes.byref(sid), ctypes.byref(nb)))
        self.scene_bytes = getattr(self, "scene_bytes", {})
        self.scene_bytes[sid.value] = nb.value
        return sid.value

    def submit(self, scene, width, height, spp, bounces=6, seed=1000, arrival_s=0.0):
        jid = ctypes.c_int()
        _check(lib().cwp_server_submit(self.h, int(scene), int(width), int(height), int(spp), int(bounces), int(seed) & 0xFFFFFFFF,
                                       float(arrival_s), ctypes.byref(jid)))
        return jid.value

    def run(self, batched=None):
        b = self.options.batched if batched is None else (1 if batched else 0)
        _check(lib().cwp_server_run(self.h, b))
        return self.stats()

    def job(self, jid):
        i = JobInfo()
        _check(lib().cwp_server_job(self.h, int(jid), ctypes.byref(i)))
        return i.as_dict()

    def image(self, jid):
        """Per-pixel radiance sums of a finished job, (height, width, 4) float32 (divide by spp)."""
        j = self.job(jid)
        out = np.zeros((j["height"], j["width"], 4), dtype=np.float32)
        _check(lib().cwp_server_image(self.h, int(jid), out.ctypes.data_as(ctypes.POINTER(ctypes.c_float)),
                                      j["width"] * j["height"]))
        return out

    def stats(self):
        s = ServeStats()
        _check(lib().cwp_server_stats(self.h, ctypes.byref(s)))
        return s.as_dict()

    # live serving: submissions while other jobs are on the GPU
    JOB_QUEUED, JOB_RUNNING, JOB_DONE = 0, 1, 2

    def start(self):
        _check(lib().cwp_server_start(self.h))

    def submit_live(self, scene, width, height, spp, bounces=6, seed=1000):
        jid = ctypes.c_int()
        _check(lib().cwp_job_submit(self.h, int(scene), int(width), int(height), int(spp), int(bounces),
                                    int(seed) & 0xFFFFFFFF, ctypes.byref(jid)))
        return jid.value

    def poll(self, jid):
        st = ctypes.c_int()
        _check(lib().cwp_job_poll(self.h, int(jid), ctypes.byref(st)))
        return st.value

    def result(self, jid, width, height):
        out = np.zeros((height, width, 4), dtype=np.float32)
        _check(lib().cwp_job_result(self.h, int(jid), out.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), width * height))
        return out

    def stop(self):
        _check(lib().cwp_server_stop(self.h))

    def clear_jobs(self):
        _check(lib().cwp_server_clear_jobs(self.h))

    def reset_jobs(self):
        _check(lib().cwp_server_reset_jobs(self.h))

    def preload(self):
        _check(lib().cwp_server_preload(self.h))

    def get_options(self):
        o = ServerOptions()
        _check(lib().cwp_server_get_options(self.h, ctypes.byref(o)))
        return o.as_dict()

    def autotune(self, variants=(4, 8), stacks=(8, 16, 32), rays_per_lane=(1, 2, 4), compacts=(0,)):
        """Runs the submitted jobs under every combination (compacts: 0 per-wave, 1 per-lane queue compaction), keeps
        the fastest; returns (trials, chosen options)."""
        def arr(x):
            return (ctypes.c_int * len(x))(*x)
        n = len(variants) * len(stacks) * len(rays_per_lane) * len(compacts)
        trials = (TuneTrial * n)()
        got, chosen = ctypes.c_int(), ServerOptions()
        _check(lib().cwp_server_autotune(self.h, arr(variants), len(variants), arr(stacks), len(stacks), arr(rays_per_lane),
                                         len(rays_per_lane), arr(compacts), len(compacts), trials, n, ctypes.byref(got),
                                         ctypes.byref(chosen)))
        self.options = chosen
        return [trials[k].as_dict() for k in range(got.value)], chosen.as_dict()

    def close(self):
        if getattr(self, "h", None):
            lib().cwp_server_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def image_read(path):
    """RGBA8 texels of a PNG / TGA / PPM / PGM file as (h, w) uint32 (r in the low byte), rows top first."""
    w, h = ctypes.c_int(), ctypes.c_int()
    _check(lib().cwp_image_read(str(path).encode(), ctypes.byref(w), ctypes.byref(h), None, 0))
    out = np.zeros((h.value, w.value), np.uint32)
    _check(lib().cwp_image_read(str(path).encode(), ctypes.byref(w), ctypes.byref(h), out.ctypes.data, out.size))
    return out


def inflate(data, capacity):
    """zlib stream -> bytes (at most capacity)."""
    src = np.frombuffer(bytes(data), np.uint8)
    out = np.zeros(max(capacity, 1), np.uint8)
    n = ctypes.c_size_t()
    _check(lib().cwp_inflate(src.ctypes.data, src.size, out.ctypes.data, capacity, ctypes.byref(n)))
    return out[:n.value].tobytes()


def texture_sample(image, texels, u, v, nearest=False):
    """The kernels' texture lookup on the host: image = (first, w, h, srgb), texels uint32; returns rgba (4,)."""
    im = (ctypes.c_int * 4)(*[int(x) for x in image])
    tx = np.ascontiguousarray(texels, np.uint32)
    out = (ctypes.c_float * 4)()
    _check(lib().cwp_texture_sample(im, tx.ctypes.data, tx.size, float(u), float(v), 1 if nearest else 0, out))
    return np.array(out[:], np.float32)


def bsdf_eval(mat, wo, wi, eta, gd_energy=1, ux1=0.5, ux2=0.5):
    """f (3,) and pdf of the full integrator's material response (the kernels' code, run on the host)."""
    m = np.ascontiguousarray(mat, dtype=np.float32).reshape(16)
    a = np.ascontiguousarray(wo, dtype=np.float32).reshape(3)
    b = np.ascontiguousarray(wi, dtype=np.float32).reshape(3)
    f = np.zeros(3, np.float32)
    p = np.zeros(1, np.float32)
    _check(lib().cwp_bsdf_eval(_fptr(m), _fptr(a), _fptr(b), eta, gd_energy, ux1, ux2, _fptr(f), _fptr(p)))
    return f, float(p[0])


def bsdf_sample(mat, wo, eta, u, u1, u2, gd_energy=1, ux1=0.5, ux2=0.5):
    """(ok, wi (3,), weight (3,), pdf) of one sample of the material response; pdf 0 marks a delta sample."""
    m = np.ascontiguousarray(mat, dtype=np.float32).reshape(16)
    a = np.ascontiguousarray(wo, dtype=np.float32).reshape(3)
    wi = np.zeros(3, np.float32)
    w = np.zeros(3, np.float32)
    p = np.zeros(1, np.float32)
    ok = ctypes.c_int(0)
    _check(lib().cwp_bsdf_sample(_fptr(m), _fptr(a), eta, gd_energy, ux1, ux2, u, u1, u2, _fptr(wi), _fptr(w), _fptr(p),
                                 ctypes.byref(ok)))
    return bool(ok.value), wi, w, float(p[0])


def bvh8_device_nodes(nodes):
    """Device encoding (n, 32) float32 of compressed BVH8 nodes given as (n, 20) float32 (CWP_ARR_NODES)."""
    n = np.ascontiguousarray(nodes, dtype=np.float32).reshape(-1, 20)
    out = np.zeros((n.shape[0], 32), dtype=np.float32)
    _check(lib().cwp_bvh8_device_nodes(n.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), n.shape[0],
                                       out.ctypes.data_as(ctypes.POINTER(ctypes.c_float))))
    return out


def bvh_build(vertices, triangles, variant=8, bins=16, leaf_max=4):
    """Binned-SAH BVH (+ CWBVH collapse for variant 8) of caller triangles. Returns a dict of numpy arrays."""
    v = np.ascontiguousarray(vertices, dtype=np.float32).reshape(-1, 3)
    t = np.ascontiguousarray(triangles, dtype=np.int32).reshape(-1, 3)
    h = ctypes.c_void_p()
    _check(lib().cwp_bvh_build(v.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), v.shape[0],
                               t.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), t.shape[0], variant, bins, leaf_max,
                               ctypes.byref(h)))
    return _bvh_result(h, variant)


def _bvh_result(h, variant):
    try:
        out = {}
        nf4 = {2: 4, 4: 8, 8: 5}[variant]
        for name, which in (("nodes", ARR_NODES), ("order", ARR_ORDER), ("woop", ARR_WOOP), ("corners", ARR_CORNERS),
                            ("tree", ARR_TREE), ("tree_box", ARR_TREE_BOX), ("prim", ARR_PRIM)):
            ptr, cnt = ctypes.c_void_p(), ctypes.c_size_t()
            _check(lib().cwp_bvh_array(h, which, ctypes.byref(ptr), ctypes.byref(cnt)))
            out[name] = _view(ptr.value, cnt.value, which, nf4)
        d, s, sah = ctypes.c_int(), ctypes.c_int(), ctypes.c_double()
        _check(lib().cwp_bvh_stats(h, ctypes.byref(d), ctypes.byref(s), ctypes.byref(sah)))
        out.update(depth=d.value, stack_need=s.value, sah=sah.value, variant=variant)
        return out
    finally:
        lib().cwp_bvh_destroy(h)


def _fptr(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


def _boxes4(a):
    a = np.ascontiguousarray(a, dtype=np.float32)
    if a.ndim == 2 and a.shape[1] == 3:
        a = np.concatenate([a, np.zeros((len(a), 1), np.float32)], 1)
    return np.ascontiguousarray(a.reshape(-1, 4))


MORTON_ADAPTIVE, MORTON_30 = 0, 30


def hploc_build(lo, hi, morton_bits=MORTON_ADAPTIVE):
    """GPU H-PLOC build over host boxes ((n, 3) or (n, 4)). Returns (nodes (2n-1, 8) float32, root, gpu_ms).
    morton_bits: MORTON_ADAPTIVE (the reference's keys) or MORTON_30 (10 bits per axis)."""
    lo, hi = _boxes4(lo), _boxes4(hi)
    n = len(lo)
    nodes = np.zeros((2 * n - 1, 8), np.float32)
    root, ms = ctypes.c_int(), ctypes.c_double()
    _check(lib().cwp_hploc_build_host_mode(_fptr(lo), _fptr(hi), n, morton_bits, _fptr(nodes), ctypes.byref(root),
                                           ctypes.byref(ms)))
    return nodes, root.value, ms.value


def cwbvh_from_hploc_gpu(hnodes, n):
    """Compressed BVH8 from H-PLOC nodes on the GPU. Returns (nodes (m, 20) float32, order (n,) int32, depth, gpu_ms)."""
    h = np.ascontiguousarray(hnodes, dtype=np.float32).reshape(-1, 8)
    assert h.shape[0] == 2 * n - 1
    nodes = np.zeros((n, 20), np.float32)
    prims = np.zeros(n, np.int32)
    m, d, ms = ctypes.c_int(), ctypes.c_int(), ctypes.c_double()
    _check(lib().cwp_cwbvh_from_hploc_host(_fptr(h), n, _fptr(nodes), ctypes.byref(m), prims.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                                           ctypes.byref(d), ctypes.byref(ms)))
    return nodes[:m.value].copy(), prims, d.value, ms.value


def bvh_from_hploc(vertices, triangles, hnodes, root, variant=8, leaf_max=4):
    """Layout (as bvh_build) from H-PLOC nodes over the mesh's triangle boxes."""
    v = np.ascontiguousarray(vertices, dtype=np.float32).reshape(-1, 3)
    t = np.ascontiguousarray(triangles, dtype=np.int32).reshape(-1, 3)
    h = np.ascontiguousarray(hnodes, dtype=np.float32)
    # the C entry point reads 2 * (2n - 1) float4 of H-PLOC nodes for the n triangles
    if h.size != (2 * t.shape[0] - 1) * 8:
        raise ValueError("bvh_from_hploc: %d floats of H-PLOC nodes for %d triangles (expected %d)"
                         % (h.size, t.shape[0], (2 * t.shape[0] - 1) * 8))
    p = ctypes.c_void_p()
    _check(lib().cwp_bvh_from_hploc(_fptr(v), v.shape[0], t.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), t.shape[0],
                                    _fptr(h), int(root), variant, leaf_max, ctypes.byref(p)))
    return _bvh_result(p, variant)


DISPLAY_MODES = {"clip": 0, "reinhard": 1, "aces": 2}


def display_buffer(accum, spp, exposure_ev=0.0, mode="clip"):
    """The display transform kernel on a host array of (n, 4) float32 sums (staged through device memory):
    (n, 4) uint8 RGBA. Mode: clip, reinhard, aces (or 0, 1, 2)."""
    a = np.ascontiguousarray(accum, np.float32).reshape(-1, 4)
    n = a.shape[0]
    m = DISPLAY_MODES[mode] if isinstance(mode, str) else int(mode)
    src = DeviceArray.from_host(a) if n else None
    dst = DeviceArray(max(n, 1) * 4)
    _check(lib().cwp_display_buffer(src.ptr if src else None, dst.ptr, n, int(spp), float(exposure_ev), m))
    return dst.to_host(np.uint8, (n, 4))


def srgb_encode8(x):
    """the PNG writer's 8-bit encode of one value (clip, sRGB OETF, x * 255 + 0.5 truncated)"""
    return lib().cwp_srgb_encode8(float(x))


def srgb_steps():
    """the 256 step positions the display kernel searches (steps[0] = -inf)"""
    t = np.zeros(256, np.float32)
    lib().cwp_srgb_steps(t.ctypes.data_as(ctypes.POINTER(ctypes.c_float)))
    return t


class DeviceArray:
    """Device memory owned by libcwp (for the device-pointer entry points without a GPU array library)."""

    def __init__(self, nbytes):
        p = ctypes.c_void_p()
        _check(lib().cwp_device_malloc(nbytes, ctypes.byref(p)))
        self.ptr, self.nbytes = p.value, nbytes

    @classmethod
    def from_host(cls, a):
        a = np.ascontiguousarray(a)
        d = cls(a.nbytes)
        _check(lib().cwp_memcpy(d.ptr, a.ctypes.data, a.nbytes, 1))
        return d

    def to_host(self, dtype, shape):
        out = np.zeros(shape, dtype)
        _check(lib().cwp_memcpy(out.ctypes.data, self.ptr, out.nbytes, 2))
        return out

    def free(self):
        if self.ptr:
            lib().cwp_device_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass
