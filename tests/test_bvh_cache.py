"""On-disk formats of the host side: built-scene cache files (--bvh_cache) and the Radiance HDR writer (--env_out).

A cached build must be indistinguishable from a fresh one: the second setup of the same scene loads the file and its
GPU layout, packed triangles, binary tree and statistics are bit-identical to a setup without the cache. Keys separate
layouts and meshes; damaged files are rebuilt, never used. These run on the CPU only (no device calls).
"""
import os

import numpy as np
import pytest

import cwp

SMALL = ["--width", "8", "--height", "8"]
GEOMETRY = [cwp.ARR_NODES, cwp.ARR_ORDER, cwp.ARR_WOOP, cwp.ARR_CORNERS, cwp.ARR_TREE, cwp.ARR_TREE_BOX, cwp.ARR_PRIM]


def _same_geometry(a, b, extra=()):
    for which in list(GEOMETRY) + list(extra):
        x, y = a.array(which), b.array(which)
        assert x.shape == y.shape, which
        np.testing.assert_array_equal(x.view(np.uint8), y.view(np.uint8), err_msg=str(which))
    for k in ("bvh_depth", "sah", "num_nodes", "stack_need", "variant"):
        assert a.info[k] == b.info[k], k


def _files(d):
    return sorted(f for f in os.listdir(d) if f.endswith(".cwpbvh"))


@pytest.mark.parametrize("variant", ["bvh2", "bvh4", "bvh8"])
def test_second_setup_loads_the_cached_build(tmp_path, variant):
    args = ["--gen", "boxes", "--variant", variant] + SMALL
    plain = cwp.Setup(args)
    assert plain.info["bvh_cache"] == -1
    first = cwp.Setup(args + ["--bvh_cache", str(tmp_path)])
    assert first.info["bvh_cache"] == 0 and len(_files(tmp_path)) == 1
    second = cwp.Setup(args + ["--bvh_cache", str(tmp_path)])
    assert second.info["bvh_cache"] == 1
    _same_geometry(plain, first)
    _same_geometry(plain, second)


def test_keys_separate_layouts_meshes_and_builder_options(tmp_path):
    c = ["--bvh_cache", str(tmp_path)] + SMALL
    cwp.Setup(["--gen", "boxes"] + c)
    assert cwp.Setup(["--gen", "boxes", "--variant", "bvh4"] + c).info["bvh_cache"] == 0
    assert cwp.Setup(["--gen", "boxes", "--gen_seed", "7"] + c).info["bvh_cache"] == 0
    assert cwp.Setup(["--gen", "boxes", "--variant", "bvh4", "--leaf", "2"] + c).info["bvh_cache"] == 0
    assert cwp.Setup(["--gen", "boxes", "--bins", "32"] + c).info["bvh_cache"] == 0
    assert len(_files(tmp_path)) == 5
    # the camera, resolution and lighting are not part of the key
    assert cwp.Setup(["--gen", "boxes", "--bvh_cache", str(tmp_path), "--width", "16", "--height", "4", "--spp", "3",
                      "--sky", "1", "1", "1"]).info["bvh_cache"] == 1


@pytest.mark.parametrize("damage", ["truncate", "flip", "append"])
def test_damaged_cache_files_are_rebuilt(tmp_path, damage):
    args = ["--gen", "terrain", "--gen_grid", "41", "--gen_boxes", "10"] + SMALL
    plain = cwp.Setup(args)
    cwp.Setup(args + ["--bvh_cache", str(tmp_path)])
    (name,) = _files(tmp_path)
    path = os.path.join(str(tmp_path), name)
    data = bytearray(open(path, "rb").read())
    if damage == "truncate":
        data = data[: len(data) // 2]
    elif damage == "flip":
        data[len(data) // 3] ^= 0x40
    else:
        data += b"\0"
    open(path, "wb").write(bytes(data))
    again = cwp.Setup(args + ["--bvh_cache", str(tmp_path)])
    assert again.info["bvh_cache"] == 0          # not trusted: built again and rewritten
    _same_geometry(plain, again)
    assert cwp.Setup(args + ["--bvh_cache", str(tmp_path)]).info["bvh_cache"] == 1


def test_two_level_scene_cache(tmp_path):
    args = ["--gen", "instances", "--gen_small", "1", "--gen_instances", "40"] + SMALL
    plain = cwp.Setup(args)
    cwp.Setup(args + ["--bvh_cache", str(tmp_path)])
    cached = cwp.Setup(args + ["--bvh_cache", str(tmp_path)])
    assert cached.info["bvh_cache"] == 1
    _same_geometry(plain, cached, extra=[cwp.ARR_MESH_ROOTS, cwp.ARR_MESH_BOXES, cwp.ARR_MESH_TRIS])
    for f in (0, 3):
        np.testing.assert_array_equal(plain.instances(f).view(np.uint8), cached.instances(f).view(np.uint8))


def test_full_scene_cache_keeps_emitter_tables(tmp_path):
    args = ["--gen", "boxes", "--gen_materials", "3", "--integrator", "full"] + SMALL
    plain = cwp.Setup(args)
    cwp.Setup(args + ["--bvh_cache", str(tmp_path)])
    cached = cwp.Setup(args + ["--bvh_cache", str(tmp_path)])
    assert cached.info["bvh_cache"] == 1
    _same_geometry(plain, cached, extra=[cwp.ARR_MATERIALS, cwp.ARR_EM_TRI, cwp.ARR_EM_CDF, cwp.ARR_TRI_LIGHT_PDF])


def test_env_out_writes_the_generated_map(tmp_path):
    """--env_out writes the environment radiance as RGBE; read back through --env it is the same map rounded down to
    8 significant bits per channel (relative error below 2^-7 of the pixel's largest channel)."""
    out = str(tmp_path / "env.hdr")
    a = cwp.Setup(["--gen", "boxes", "--gen_env", "64,32,1.5,4", "--env_out", out] + SMALL)
    with open(out, "rb") as f:
        head = f.read(64)
    assert head.startswith(b"#?RADIANCE\nFORMAT=32-bit_rle_rgbe\n\n-Y 32 +X 64\n")
    b = cwp.Setup(["--gen", "boxes", "--env", out] + SMALL)
    assert (b.info["env_w"], b.info["env_h"]) == (64, 32)
    ta, tb = a.array(cwp.ARR_ENV_TEX)[:, :3], b.array(cwp.ARR_ENV_TEX)[:, :3]
    m = ta.max(axis=1, keepdims=True)
    assert np.all(tb <= ta + 1e-30)
    assert np.all(ta - tb <= m * 2.0 ** -7)
    with pytest.raises(cwp.CwpError, match="env_out"):
        cwp.Setup(["--gen", "boxes", "--env_out", out] + SMALL)
