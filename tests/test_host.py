"""Host-side tests of libcwp.so that need no GPU: the C ABI exports, option parsing and error codes, the reference's
file formats (its examples/ fixtures), and bit-exact agreement of the product's BVH build and triangle packing with the
oracle's independent restatement."""
import ctypes
import json
import os
import re

import numpy as np
import pytest

import cwp
import pyoracle as po

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "cwp.h")).read()
    return sorted(set(re.findall(r"^\s*(?:const\s+)?\w+\**\s+\**(cwp_\w+)\s*\(", text, re.M)))


def test_library_exports_every_declared_symbol():
    L = ctypes.CDLL(cwp.LIB_PATH)
    names = declared_symbols()
    assert len(names) >= 20
    for n in names:
        assert hasattr(L, n), n
    assert set(names) == set(cwp.SIGNATURES)   # the binding covers the whole header


def test_version_and_error_string():
    assert cwp.lib().cwp_version() == 1
    with pytest.raises(cwp.CwpError, match="unknown arg"):
        cwp.Setup(["--no_such_flag"])
    with pytest.raises(cwp.CwpError, match="needs 1 value"):
        cwp.Setup(["--gen", "boxes", "--spp"])
    with pytest.raises(cwp.CwpError, match="--variant"):
        cwp.Setup(["--gen", "boxes", "--variant", "bvh3"])
    with pytest.raises(cwp.CwpError, match="failed to load"):
        cwp.Setup(["--mesh", "/nonexistent.obj", "--hfov", "40"])
    with pytest.raises(cwp.CwpError, match="no scene"):
        cwp.Setup([])
    with pytest.raises(cwp.CwpError, match="--fuse"):
        cwp.Setup(["--gen", "boxes", "--fuse", "2"])
    with pytest.raises(cwp.CwpError, match="--persistent"):
        cwp.Setup(["--gen", "boxes", "--persistent", "3"])
    with pytest.raises(cwp.CwpError, match="--sort"):
        cwp.Setup(["--gen", "boxes", "--sort", "3"])
    with pytest.raises(cwp.CwpError, match="experimental"):
        cwp.Setup(["--gen", "boxes", "--devloop", "1"])
    with pytest.raises(cwp.CwpError, match="--morton_bits"):
        cwp.Setup(["--gen", "boxes", "--morton_bits", "12"])
    with pytest.raises(cwp.CwpError, match="--tlas"):
        cwp.Setup(["--gen", "boxes", "--tlas", "bvh4"])
    for bad in ("1", "65", "-2"):
        with pytest.raises(cwp.CwpError, match="--tail"):
            cwp.Setup(["--gen", "boxes", "--tail", bad])
    cwp.Setup(["--gen", "boxes", "--sort", "1", "--devloop", "0", "--morton_bits", "30", "--tlas", "bvh2"])
    assert cwp.Options.default().sort == 0
    assert cwp.Options.default().tail_lanes == -1   # automatic: on where the schedule supports it


def test_reference_example_fixtures(golden_dir):
    s = cwp.Setup(["--mesh", os.path.join(golden_dir, "boxes.obj"), "--camera", os.path.join(golden_dir, "camera.json"),
                   "--lighting", os.path.join(golden_dir, "lighting.json"), "--width", "960", "--height", "540"])
    i = s.info
    assert i["num_tris"] == 62 and i["num_verts"] == 124
    kd = s.array(cwp.ARR_ALBEDO)
    assert len(kd) == 7   # default + 6 MTL materials
    np.testing.assert_allclose(kd[2, :3], [0.8, 0.35, 0.22])   # terracotta
    p = s.params()
    assert p.fx == pytest.approx(1158.823) and p.fy == pytest.approx(1158.823)   # width/height match the file
    assert p.pos[:] == pytest.approx([9.0, -10.5, 5.5])
    sun = np.array([0.608749, -0.42625, 0.669131]); sun /= np.linalg.norm(sun)
    assert np.array(p.sun_to[:]) == pytest.approx(sun, abs=1e-6)
    assert p.sun_e[:] == pytest.approx([3.0] * 3) and p.sky[:] == pytest.approx([0.3, 0.38, 0.5])
    assert p.sun_cos == pytest.approx(np.cos(0.5 * 0.00918), abs=1e-7)
    mats = s.array(cwp.ARR_MATERIAL)
    assert set(mats.tolist()) <= set(range(1, 7))
    # other resolution: hfov from the file is used instead of fx
    s2 = cwp.Setup(["--mesh", os.path.join(golden_dir, "boxes.obj"), "--camera", os.path.join(golden_dir, "camera.json"),
                    "--width", "320", "--height", "180"])
    assert s2.params().fx == pytest.approx(160 / np.tan(np.radians(22.5)), rel=1e-6)


def test_ply_loader_roundtrip(tmp_path):
    v = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0]], np.float32)
    hdr = "ply\nformat binary_little_endian 1.0\nelement vertex 4\nproperty float x\nproperty float y\nproperty float z\n" \
          "element face 1\nproperty list uchar int vertex_indices\nproperty int material_index\nend_header\n"
    body = v.tobytes() + bytes([4]) + np.array([0, 1, 2, 3], np.int32).tobytes() + np.array([2], np.int32).tobytes()
    path = tmp_path / "quad.ply"
    path.write_bytes(hdr.encode() + body)
    s = cwp.Setup(["--mesh", str(path), "--c2w", "1,0,0,0.5,0,1,0,0.5,0,0,1,2,0,0,0,1"])
    assert s.info["num_tris"] == 2
    np.testing.assert_array_equal(s.array(cwp.ARR_TRIANGLES), [[0, 1, 2], [0, 2, 3]])
    np.testing.assert_array_equal(s.array(cwp.ARR_MATERIAL), [2, 2])
    assert s.info["num_materials"] == 3


@pytest.mark.parametrize("gen,variant,bins", [("boxes", 8, 16), ("boxes", 2, 32), ("boxes", 8, 32), ("terrain", 8, 16),
                                              ("boxes", 4, 16), ("terrain", 4, 16)])
def test_product_bvh_matches_oracle_bit_exact(gen, variant, bins):
    args = ["--gen", gen, "--variant", "bvh%d" % variant, "--bins", str(bins), "--width", "8", "--height", "8"]
    if gen == "terrain":
        args += ["--gen_grid", "121", "--gen_boxes", "50"]
    s = cwp.Setup(args)
    v, t = s.array(cwp.ARR_VERTICES), s.array(cwp.ARR_TRIANGLES)
    b = po.build(v, t, variant, bins, 4)
    np.testing.assert_array_equal(s.array(cwp.ARR_NODES).view(np.uint32), b["nodes"].view(np.uint32))
    np.testing.assert_array_equal(s.array(cwp.ARR_ORDER), b["order"])
    np.testing.assert_array_equal(s.array(cwp.ARR_WOOP).view(np.uint32), b["woop"].view(np.uint32))
    np.testing.assert_array_equal(s.array(cwp.ARR_TREE), b["tree"])
    np.testing.assert_array_equal(s.array(cwp.ARR_TREE_BOX).view(np.uint32), b["tree_box"].view(np.uint32))
    assert s.info["bvh_depth"] == b["depth"]


@pytest.mark.parametrize("gen", ["boxes", "terrain"])
def test_bvh8_device_encoding_is_exact(gen):
    """The device encoding of BVH8 nodes carries the 80-byte node's fields exactly; the top byte of the child base is
    the mask of the slots whose metadata has the leaf bit."""
    args = ["--gen", gen, "--width", "8", "--height", "8"] + (["--gen_grid", "61", "--gen_boxes", "20"] if gen == "terrain" else [])
    n80 = cwp.Setup(args).array(cwp.ARR_NODES).reshape(-1, 20)
    dev = cwp.bvh8_device_nodes(n80)
    u80, ud = n80.view(np.uint32), dev.view(np.uint32)
    np.testing.assert_array_equal(ud[:, 0:4], u80[:, 0:4])
    np.testing.assert_array_equal(ud[:, 5:8], u80[:, 5:8])
    np.testing.assert_array_equal(ud[:, 4] & 0xFFFFFF, u80[:, 4])
    meta = n80[:, 6:8].copy().view(np.uint8).reshape(-1, 8)
    leaf = ((meta >> 7) & 1).astype(np.uint32) @ (1 << np.arange(8, dtype=np.uint32))
    np.testing.assert_array_equal(ud[:, 4] >> 24, leaf)
    assert leaf.any() and (u80[:, 3] >> 24).any()
    np.testing.assert_array_equal(ud[:, 8:20], u80[:, 8:20])
    assert not ud[:, 20:].any()
    bad = n80.copy()
    bad[0, 4] = np.array([1 << 24], dtype=np.int32).view(np.float32)[0]
    with pytest.raises(cwp.CwpError, match="24-bit"):
        cwp.bvh8_device_nodes(bad)


def test_bvh_build_api_matches_oracle():
    rng = np.random.default_rng(9)
    c = rng.random((700, 3)).astype(np.float32)
    v = (c[:, None, :] + rng.normal(scale=0.03, size=(700, 3, 3))).reshape(-1, 3).astype(np.float32)
    t = np.arange(2100, dtype=np.int32).reshape(-1, 3)
    for variant in (2, 4, 8):
        p = cwp.bvh_build(v, t, variant, 16, 4)
        o = po.build(v, t, variant, 16, 4)
        np.testing.assert_array_equal(p["nodes"].view(np.uint32), o["nodes"].view(np.uint32))
        np.testing.assert_array_equal(p["order"], o["order"])
        np.testing.assert_array_equal(p["corners"][:, :, :3], o["corners"][:, :, :3])
    with pytest.raises(cwp.CwpError, match="out of range"):
        cwp.bvh_build(v, t + 5000, 8)


def test_params_filter_table_matches_oracle():
    s = cwp.Setup(["--gen", "boxes", "--width", "8", "--height", "8"])
    np.testing.assert_array_equal(np.array(s.params().ftab[:], np.float32), po.blackman_harris_table(64, 1.5))


def test_scene_generators_are_deterministic_and_sized():
    a = cwp.Setup(["--gen", "terrain", "--width", "8", "--height", "8"])
    assert a.info["num_tris"] == 2 * 241 * 241 + 200 * 12 == 118562
    b = cwp.Setup(["--gen", "terrain", "--width", "8", "--height", "8"])
    np.testing.assert_array_equal(a.array(cwp.ARR_VERTICES), b.array(cwp.ARR_VERTICES))
    c = cwp.Setup(["--gen", "terrain", "--gen_seed", "2", "--width", "8", "--height", "8"])
    assert not np.array_equal(a.array(cwp.ARR_VERTICES), c.array(cwp.ARR_VERTICES))
    d = cwp.Setup(["--gen", "boxes", "--width", "8", "--height", "8"])
    assert d.info["num_tris"] == 1024 * 12 + 2


def test_cli_reports_errors_without_gpu(tmp_path):
    import subprocess
    exe = os.path.join(cwp.PKG, "pt")
    r = subprocess.run([exe, "--bogus"], capture_output=True, text=True)
    assert r.returncode == 1 and "unknown arg" in r.stderr
    r = subprocess.run([exe, "--gen", "boxes", "--spp", "10", "--batch", "4", "--width", "8", "--height", "8"],
                       capture_output=True, text=True)
    assert r.returncode == 1 and "multiple of batch" in r.stderr


@pytest.mark.parametrize("gen,variant", [("boxes", 2), ("boxes", 8), ("terrain", 8), ("boxes", 4)])
def test_scene_check_accepts_built_scenes_and_rejects_bad_indices(gen, variant):
    s = cwp.Setup(["--gen", gen, "--variant", "bvh%d" % variant, "--width", "8", "--height", "8"])
    nodes = s.array(cwp.ARR_NODES).copy()
    corners = s.array(cwp.ARR_CORNERS).copy()
    albedo = s.array(cwp.ARR_ALBEDO).copy()
    woop = s.array(cwp.ARR_WOOP)
    need = cwp.scene_check(variant, nodes, corners, albedo, woop)
    assert need == s.info["stack_need"]
    ints = nodes.view(np.int32)
    nn, nt = len(nodes), len(corners)

    def rejected(n=nodes, c=corners, a=albedo, match=None):
        with pytest.raises(cwp.CwpError, match=match):
            cwp.scene_check(variant, n, c, a, woop)

    if variant == 4:
        internal = np.nonzero((ints[:, 28] == 0) & (ints[:, 24] > 0))[0]
        leaf = np.nonzero(ints[:, 28] > 0)[0]
        bad = ints.copy(); bad[internal[0], 24] = nn; rejected(bad.view(np.float32), match="child index")
        bad = ints.copy(); bad[internal[-1], 24] = 0; rejected(bad.view(np.float32), match="parent")
        bad = ints.copy(); bad[leaf[0], 24] = nt - 1; bad[leaf[0], 28] = 2; rejected(bad.view(np.float32), match="triangle range")
    elif variant == 2:
        internal = np.nonzero((ints[:, 14] == 0) & (ints[:, 12] > 0))[0]
        leaf = np.nonzero(ints[:, 14] > 0)[0]
        bad = ints.copy(); bad[internal[0], 12] = nn; rejected(bad.view(np.float32), match="child index")
        bad = ints.copy(); bad[internal[1], 12] = ints[internal[0], 12]; rejected(bad.view(np.float32), match="more than one parent")
        bad = ints.copy(); bad[internal[-1], 12] = 0; rejected(bad.view(np.float32), match="parent")
        bad = ints.copy(); bad[leaf[0], 12] = nt - 1; bad[leaf[0], 14] = 2; rejected(bad.view(np.float32), match="triangle range")
        bad = ints.copy(); bad[internal[0], 12] = -1; rejected(bad.view(np.float32), match="empty slot")
    else:
        ew = ints[:, 3].view(np.uint32)
        with_kids = np.nonzero(ew >> 24)[0]
        bad = ints.copy(); bad[with_kids[-1], 4] = 0; rejected(bad.view(np.float32), match="parent")
        bad = ints.copy(); bad[with_kids[0], 4] = nn; rejected(bad.view(np.float32), match="outside the node array")
        bad = ints.copy(); bad.view(np.uint32)[with_kids[0], 3] ^= np.uint32(0xff << 24)
        rejected(bad.view(np.float32), match="mask")
        meta = ints[:, 6:8].view(np.uint8).reshape(nn, 8)
        with_leaf = np.nonzero((meta & 0x80).any(axis=1))[0]
        bad = ints.copy(); bad[with_leaf[0], 5] = nt; rejected(bad.view(np.float32), match="triangle range")
    bad_c = corners.copy(); bad_c.view(np.int32)[5, 1, 3] = len(albedo); rejected(c=bad_c, match="material")
    rejected(c=corners[:0], match="required")


def test_scene_check_accepts_the_reference_edge_form():
    s = cwp.Setup(["--gen", "boxes", "--width", "8", "--height", "8"])
    c = s.array(cwp.ARR_CORNERS)
    edges = c.copy()
    edges[:, 1, :3] = c[:, 1, :3] - c[:, 0, :3]
    edges[:, 2, :3] = c[:, 2, :3] - c[:, 0, :3]
    need = cwp.scene_check(8, s.array(cwp.ARR_NODES), edges, s.array(cwp.ARR_ALBEDO), tri_form=1)
    assert need == s.info["stack_need"]
    with pytest.raises(cwp.CwpError, match="tri_form"):
        cwp.scene_check(8, s.array(cwp.ARR_NODES), edges, s.array(cwp.ARR_ALBEDO), tri_form=2)


@pytest.mark.parametrize("variant", [2, 8])
def test_bvh_from_hploc_matches_oracle_bit_exact(variant):
    """The product's SAH leaf collapse of H-PLOC nodes and the layouts after it equal the oracle's, given the same
    H-PLOC nodes (the oracle's H-PLOC restatement here; the GPU build is compared in the GPU tests)."""
    s = cwp.Setup(["--gen", "terrain", "--gen_grid", "61", "--gen_boxes", "20", "--width", "8", "--height", "8"])
    v, t = s.array(cwp.ARR_VERTICES), s.array(cwp.ARR_TRIANGLES)
    hn, root, _ = po.hp_build(*po.tri_boxes(v, t))
    p = cwp.bvh_from_hploc(v, t, hn, root, variant, 4)
    o = po.build(v, t, variant, leaf_max=4, builder="hploc", hnodes=hn, root=root)
    np.testing.assert_array_equal(p["tree"], o["tree"])
    np.testing.assert_array_equal(p["tree_box"].view(np.uint32), o["tree_box"].view(np.uint32))
    np.testing.assert_array_equal(p["nodes"].view(np.uint32), o["nodes"].view(np.uint32))
    np.testing.assert_array_equal(p["order"], o["order"])
    np.testing.assert_array_equal(p["woop"].view(np.uint32), o["woop"].view(np.uint32))
    bad = hn.copy()
    bad.view(np.int32)[len(t), 3] = 2 * len(t) - 2   # a child id not below its parent
    with pytest.raises(cwp.CwpError, match="smaller than the parent"):
        cwp.bvh_from_hploc(v, t, bad, root, variant, 4)


def test_scene_check_accepts_hploc_bvh2_with_root_last():
    """The H-PLOC BVH2 layout stores children before parents with the root last: accepted with its root index."""
    s = cwp.Setup(["--gen", "boxes", "--width", "8", "--height", "8"])
    v, t = s.array(cwp.ARR_VERTICES), s.array(cwp.ARR_TRIANGLES)
    hn, root, _ = po.hp_build(*po.tri_boxes(v, t))
    n = len(t)
    bvh2 = po.hp_to_bvh2(hn, n)
    corners = np.zeros((n, 3, 4), np.float32)   # the layout's leaves name primitives: corners in mesh order
    corners[:, :, :3] = v[t]
    need = cwp.scene_check(2, bvh2, corners, s.array(cwp.ARR_ALBEDO), root=n - 2)
    assert need >= 2
    with pytest.raises(cwp.CwpError, match="root has a parent"):
        cwp.scene_check(2, bvh2, corners, s.array(cwp.ARR_ALBEDO), root=0)


def _rows(rec, which):
    r = rec[:, 0:12] if which == "inv" else rec[:, 12:24]
    return r.reshape(-1, 3, 4).astype(np.float64)


def test_scene_file_instances_and_transforms(tmp_path, golden_dir):
    """Two-level scene file (reference format): explicit matrix, parametric rotation / scale / per-frame motion,
    per-frame matrices. Records: fwd = the transform, inv = its inverse, det sign of a mirrored instance, BLAS roots."""
    import shutil
    shutil.copy(os.path.join(golden_dir, "boxes.obj"), tmp_path / "boxes.obj")
    shutil.copy(os.path.join(golden_dir, "boxes.mtl"), tmp_path / "boxes.mtl")
    scene = {"frames": 4,
             "meshes": [{"file": "boxes.obj"}, {"file": "boxes.obj"}],
             "instances": [
                 {"mesh": 0, "transform": [1, 0, 0, 10, 0, 1, 0, 0, 0, 0, 1, 0]},
                 {"mesh": 1, "translate": [0, 5, 0], "axis": [0, 0, 1], "angle_deg": 90, "scale": 2,
                  "translate_per_frame": [1, 0, 0], "angle_deg_per_frame": 90},
                 {"mesh": 0, "scale": [-1, 1, 1]},
                 {"mesh": 1, "transforms": [[1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0], [1, 0, 0, 3, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1]]}]}
    (tmp_path / "scene.json").write_text(json.dumps(scene))
    s = cwp.Setup(["--scene", str(tmp_path / "scene.json"), "--hfov", "40", "--c2w",
                   "1,0,0,0,0,0,-1,-30,0,1,0,5,0,0,0,1", "--width", "16", "--height", "16"])
    i = s.info
    assert i["num_meshes"] == 2 and i["num_instances"] == 4 and i["frames"] == 4
    roots = s.array(cwp.ARR_MESH_ROOTS)
    assert len(roots) == 2 and roots[0] != roots[1]
    r0 = s.instances(0)
    ints = r0.view(np.int32)
    np.testing.assert_array_equal(ints[:, 25], [0, 1, 0, 1])
    np.testing.assert_array_equal(ints[:, 24], roots[ints[:, 25]])
    fwd, inv = _rows(r0, "fwd"), _rows(r0, "inv")
    assert fwd[0].tolist() == [[1, 0, 0, 10], [0, 1, 0, 0], [0, 0, 1, 0]]
    # 90 degrees about z, scale 2, translate (0, 5, 0): x axis -> 2 y
    np.testing.assert_allclose(fwd[1][:, :3] @ [1, 0, 0], [0, 2, 0], atol=1e-6)
    np.testing.assert_allclose(fwd[1][:, 3], [0, 5, 0])
    for k in range(4):   # inverse rows undo the forward rows
        F = np.vstack([fwd[k], [0, 0, 0, 1]]); Iv = np.vstack([inv[k], [0, 0, 0, 1]])
        np.testing.assert_allclose(Iv @ F, np.eye(4), atol=1e-5)
    np.testing.assert_array_equal(r0[:, 26], [1, 1, -1, 1])   # mirrored instance: det sign -1
    r2 = s.instances(2)
    f2 = _rows(r2, "fwd")
    np.testing.assert_allclose(f2[1][:, 3], [2, 5, 0])                       # 2 frames of (1, 0, 0) motion
    np.testing.assert_allclose(f2[1][:, :3] @ [1, 0, 0], [0, -2, 0], atol=1e-6)   # 270 degrees
    assert f2[3][0, 3] == 3 and _rows(s.instances(9), "fwd")[3][0, 3] == 3   # per-frame list clamps at its end
    with pytest.raises(cwp.CwpError, match="out of range"):
        bad = dict(scene, instances=[{"mesh": 5}])
        (tmp_path / "bad.json").write_text(json.dumps(bad))
        cwp.Setup(["--scene", str(tmp_path / "bad.json"), "--hfov", "40", "--c2w", "1,0,0,0,0,1,0,0,0,0,1,5,0,0,0,1"])


def test_generated_instance_scene_and_oracle_two_level_trace_match_flattened():
    """The generated two-level scene (5 meshes + ground, random rigid transforms per frame): the oracle's two-level
    traversal agrees with the flattened scene (every instance transformed into world space) on hit / miss and
    distance, and the transforms change between frames but not between runs."""
    args = ["--gen", "instances", "--gen_small", "1", "--gen_instances", "60", "--width", "48", "--height", "32"]
    s = cwp.Setup(args)
    assert s.info["num_meshes"] == 6 and s.info["num_instances"] == 61
    r0, r0b, r1 = s.instances(0), cwp.Setup(args).instances(0), s.instances(1)
    np.testing.assert_array_equal(r0, r0b)
    assert not np.array_equal(r0[1:, 12:24], r1[1:, 12:24])
    fwd = _rows(r0, "fwd")[1:, :, :3]
    np.testing.assert_allclose(np.einsum("nij,nkj->nik", fwd, fwd), np.broadcast_to(np.eye(3), fwd.shape), atol=1e-5)   # rigid
    sc = po.Scene(8, 1, s.array(cwp.ARR_NODES), s.array(cwp.ARR_WOOP), s.array(cwp.ARR_CORNERS), s.array(cwp.ARR_ALBEDO))
    sc.set_instances(r0, s.array(cwp.ARR_MESH_BOXES))
    p = s.params()
    rays = []
    for k in range(3000):
        o, d = po.camera_ray(p, (k * 7) % (p.w * p.h), k)
        rays.append(np.concatenate([o, [0], d, [1e30]]))
    rays = np.array(rays, np.float32)
    h, ids = sc.trace_inst(rays)
    f = cwp.Setup(args + ["--flatten", "1"])
    fs, _ = po.scene_from_setup(f)
    hf = fs.trace(rays)
    gi, fi = h[:, 3].view(np.int32) >= 0, hf[:, 3].view(np.int32) >= 0
    assert gi.mean() > 0.5 and (gi == fi).mean() >= 0.999
    both = gi & fi
    assert np.abs(h[both, 0] - hf[both, 0]).max() <= 1e-4 * np.abs(hf[both, 0]).max()
    assert (ids[gi] >= 0).all() and (ids[~gi] == -1).all()


def test_server_needs_a_device():
    """The server's options carry the documented defaults; without a HIP device creation fails with a device error
    (never a silent host fallback)."""
    o = cwp.ServerOptions.default()
    assert (o.variant, o.stack_lds, o.bounces, o.max_jobs, o.rays_per_lane, o.use_graph, o.batched) == (8, 8, 8, 256, 1, 1, 1)
    assert o.max_paths == 1 << 26
    try:
        cwp.device_info()
        pytest.skip("a HIP device is present")
    except cwp.CwpError:
        pass
    with pytest.raises(cwp.CwpError, match="cwp error 3"):
        cwp.Server(o)


@pytest.mark.parametrize("variant,tail,levels", [(8, 1, 2), (8, 65, 2), (8, -2, 2), (4, 12, 2), (2, 2, 2), (8, 12, 3),
                                                 (8, -1, 0)])
def test_server_hand_off_option_errors(variant, tail, levels):
    """The batch hand-off threshold is -1 (automatic), 0 or 2..64 lanes, with 1 or 2 levels, and resumes BVH8
    traversals only; all are argument errors found before any device call."""
    o = cwp.ServerOptions.default()
    assert (o.tail_lanes, o.tail_levels) == (-1, 2)
    with pytest.raises(cwp.CwpError, match="cwp error 1.*(tail|BVH8)"):
        cwp.Server(variant=variant, tail_lanes=tail, tail_levels=levels)
