"""The pt batch CLI (csrc/app/pt.cpp) end to end on the GPU: one JSON line per timed frame with every key of the
reference's line (src/app/pt.cpp:35-41), the PNG / PFM outputs, the image equal to the library's render of the same
setup, and the frame loop of a two-level scene (per-frame instances, TLAS rebuilt on the GPU)."""
import json
import os
import subprocess

import numpy as np
import pytest

import cwp

pytestmark = pytest.mark.gpu

PT = os.path.join(os.path.dirname(cwp.LIB_PATH), "pt")
# the keys of the reference's per-frame JSON line
REFERENCE_KEYS = ("rep variant tri persistent devloop sort fuse leaf batch minb scene tris bvh_nodes bvh_bytes bvh_depth "
                  "bvh_build_s w h spp bounces rays_closest rays_shadow rays_total t_trace_ms t_frame_ms mrays_s_trace "
                  "mrays_s_e2e avg_path_len vram_MB pt_blocks gpu filter smooth seed integrator lens_r focus_d "
                  "env_sampling textures builder collapse bvh_tree_ms bvh_post_ms bvh_sah graph instances flatten frame "
                  "tlas_ms").split()


def run_pt(args):
    print("case pt", " ".join(args), flush=True)
    r = subprocess.run([PT] + args, stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    return [json.loads(line) for line in r.stdout.splitlines() if line.startswith("{")]


def read_pfm(path):
    with open(path, "rb") as f:
        assert f.readline().strip() == b"PF"
        w, h = (int(x) for x in f.readline().split())
        assert float(f.readline()) < 0   # little endian
        img = np.frombuffer(f.read(), dtype="<f4").reshape(h, w, 3)
    return img[::-1]   # rows are stored bottom to top


def test_pt_frame_json_and_image(tmp_path):
    scene = ["--gen", "boxes", "--width", "64", "--height", "48", "--spp", "8", "--batch", "8", "--bounces", "3"]
    lines = run_pt(scene + ["--reps", "2", "--warm", "8", "--out", str(tmp_path / "img")])
    assert len(lines) == 2 and [j["rep"] for j in lines] == [0, 1]
    for j in lines:
        missing = [k for k in REFERENCE_KEYS if k not in j]
        assert not missing, missing
        assert j["rays_total"] == j["rays_closest"] + j["rays_shadow"] > 0
        assert j["w"] == 64 and j["h"] == 48 and j["spp"] == 8 and j["variant"] == "bvh8" and j["gpu"]
    assert lines[0]["rays_closest"] == lines[1]["rays_closest"]
    assert os.path.getsize(tmp_path / "img.png") > 0
    img = read_pfm(tmp_path / "img.pfm")
    s = cwp.Setup(scene)
    r = cwp.Renderer(s)
    r.render(8)
    ref = r.accum()[..., :3] / 8
    r.close()
    np.testing.assert_array_equal(img, ref.astype(np.float32))


def test_pt_two_level_frames(tmp_path):
    lines = run_pt(["--gen", "instances", "--gen_small", "1", "--gen_instances", "40", "--frames", "3", "--width", "48",
                    "--height", "32", "--spp", "4", "--batch", "4", "--bounces", "3", "--reps", "1", "--warm", "0",
                    "--out", str(tmp_path / "inst")])
    assert [j["frame"] for j in lines] == [0, 1, 2]
    assert all(j["instances"] >= 40 and j["tlas_ms"] > 0 for j in lines)


def test_pt_rejects_bad_options():
    r = subprocess.run([PT, "--gen", "boxes", "--spp", "6", "--batch", "4"], stdout=subprocess.PIPE,
                       stderr=subprocess.PIPE, text=True, timeout=60)
    assert r.returncode != 0 and "multiple of batch" in r.stderr
