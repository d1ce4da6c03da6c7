"""The viewer's display transform kernel and the headless viewer on the GPU: the kernel against the oracle's
restatement (src/app/display.cu:9-33) byte for byte over every mode, exposures, sample counts and special values;
the display of a render against the oracle and against the PNG that pt writes; the viewer's progressive frames
against pt's image at the same total sample count (src/app/viewer.cpp:16-17: N frames of P passes equal the batch
renderer at N * P * batch samples); and its runtime-update round (--test_updates)."""
import json
import os
import subprocess

import numpy as np
import pytest

import cwp
import pyoracle as po

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(cwp.LIB_PATH)
PT, VIEWER = os.path.join(HERE, "pt"), os.path.join(HERE, "viewer")
SCENE = ["--gen", "boxes", "--width", "64", "--height", "48", "--bounces", "3"]


def read_pfm(path):
    with open(path, "rb") as f:
        assert f.readline().strip() == b"PF"
        w, h = (int(x) for x in f.readline().split())
        assert float(f.readline()) < 0
        img = np.frombuffer(f.read(), dtype="<f4").reshape(h, w, 3)
    return img[::-1]


def png_rgba(path):
    u = cwp.image_read(str(path))
    return u.view(np.uint8).reshape(u.shape + (4,))


def run(exe, args):
    print("case", os.path.basename(exe), " ".join(args), flush=True)
    r = subprocess.run([exe] + args, stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    return [json.loads(line) for line in r.stdout.splitlines() if line.startswith("{")]


def special_sums(rng, n):
    a = rng.uniform(-1.0, 4.0, (n, 4)).astype(np.float32)
    a[::97] *= 1e30
    a[1::89] = np.float32(1e-38)
    a[2::101, 0] = np.nan
    a[3::103, 1] = np.inf
    a[4::107, 2] = -np.inf
    a[5::109] = 0.0
    return a


@pytest.mark.parametrize("mode", ["clip", "reinhard", "aces"])
def test_display_kernel_matches_the_oracle(mode):
    rng = np.random.default_rng(11)
    a = special_sums(rng, 100003)   # not a multiple of the 256-thread workgroup
    m = cwp.DISPLAY_MODES[mode]
    for spp in (0, 1, 7, 256):
        for ev in (0.0, 1.5, -2.0):
            print("case display", mode, spp, ev, flush=True)
            got = cwp.display_buffer(a, spp, ev, mode)
            want = po.display(a, spp, ev, m)
            assert np.array_equal(got, want), (spp, ev, int(np.sum(np.any(got != want, 1))))


def test_display_kernel_argument_errors():
    a = np.zeros((4, 4), np.float32)
    with pytest.raises(cwp.CwpError):
        cwp.display_buffer(a, 1, 0.0, 3)
    with pytest.raises(cwp.CwpError):
        cwp.display_buffer(a, 1, float("inf"), 0)
    with pytest.raises(cwp.CwpError):
        cwp.display_buffer(a, -1, 0.0, 0)
    assert cwp.display_buffer(np.zeros((0, 4), np.float32), 1).shape == (0, 4)


def test_display_of_a_render_equals_the_oracle_and_the_png(tmp_path):
    run(PT, SCENE + ["--spp", "8", "--batch", "4", "--warm", "0", "--out", str(tmp_path / "img")])
    s = cwp.Setup(SCENE + ["--spp", "8", "--batch", "4"])
    r = cwp.Renderer(s)
    assert r.accum_device()
    r.render(8)
    acc = r.accum()
    out = cwp.DeviceArray(64 * 48 * 4)
    for mode, ev in (("clip", 0.0), ("reinhard", 0.5), ("aces", -1.0)):
        d = r.display(8, ev, mode, out=out)
        assert np.array_equal(d.reshape(-1, 4), po.display(acc.reshape(-1, 4), 8, ev, cwp.DISPLAY_MODES[mode]))
    d = r.display(8)
    r.close()
    assert np.array_equal(d, png_rgba(tmp_path / "img.png"))


def test_viewer_frames_equal_the_batch_renderer(tmp_path):
    lines = run(VIEWER, SCENE + ["--batch", "2", "--frames", "2", "--passes", "2", "--screenshot", str(tmp_path / "v.png"),
                                 "--screenshot_pfm", str(tmp_path / "v.pfm")])
    assert [(j["frame"], j["spp"]) for j in lines] == [(0, 4), (1, 8)]
    run(PT, SCENE + ["--spp", "8", "--batch", "2", "--warm", "0", "--out", str(tmp_path / "p")])
    assert np.array_equal(read_pfm(tmp_path / "v.pfm"), read_pfm(tmp_path / "p.pfm"))
    assert np.array_equal(png_rgba(tmp_path / "v.png"), png_rgba(tmp_path / "p.png"))


def test_viewer_tone_curve_and_exposure(tmp_path):
    run(VIEWER, SCENE + ["--batch", "2", "--frames", "1", "--passes", "3", "--tonemap", "aces", "--exposure", "1",
                         "--screenshot", str(tmp_path / "a.png")])
    s = cwp.Setup(SCENE + ["--batch", "2"])
    r = cwp.Renderer(s)
    for p in range(3):
        r.render(2, sample_base=2 * p, clear=p == 0)
    acc = r.accum()
    r.close()
    want = po.display(acc.reshape(-1, 4), 6, 1.0, 2).reshape(48, 64, 4)
    assert np.array_equal(png_rgba(tmp_path / "a.png"), want)


@pytest.mark.parametrize("integrator", ["legacy", "full"])
def test_viewer_runtime_update_round(tmp_path, integrator):
    extra = ["--integrator", "full", "--gen_env", "64,32,1.0", "--gen_materials", "3"] if integrator == "full" else []
    lines = run(VIEWER, SCENE + extra + ["--batch", "2", "--frames", "2", "--test_updates", "1",
                                         "--screenshot_pfm", str(tmp_path / "u.pfm"), "--screenshot", str(tmp_path / "u.png")])
    assert len(lines) == 4 and [j.get("updated", False) for j in lines] == [False, False, True, True]
    a, b = read_pfm(tmp_path / "u.pfm"), read_pfm(tmp_path / "u_upd.pfm")
    assert np.all(np.isfinite(a)) and np.all(np.isfinite(b))
    assert a.mean() > 0 and b.mean() > 0
    assert np.mean(np.abs(a - b)) > 1e-3 * a.mean()   # camera, sun and material 0 changed
    assert os.path.getsize(tmp_path / "u_upd.png") > 0
