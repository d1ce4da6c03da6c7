"""GPU parity tests of the HIP kernels against the oracle, through the C ABI (libcwp.so).

Bars: nearest-hit triangle ids equal on >= 99.99% of rays with t within 1e-5 relative; images >= 50 dB PSNR against
the oracle (sRGB, clip [0,1], as the reference's tools/psnr.py); integer counters (rays traced) equal within 0.1%;
closed forms exact. The oracle builds its own BVH from the same mesh (independently of the product's builder).
"""
import math
import os

import numpy as np
import pytest

import cwp
import pyoracle as po

pytestmark = pytest.mark.gpu


def psnr(a, b):
    def srgb(x):
        x = np.clip(x, 0, 1)
        return np.where(x <= 0.0031308, 12.92 * x, 1.055 * np.power(x, 1 / 2.4) - 0.055)
    mse = float(np.mean((srgb(a.astype(np.float64)) - srgb(b.astype(np.float64))) ** 2))
    return 99.0 if mse == 0 else 10 * math.log10(1.0 / mse)


def scene_rays(setup, n, seed):
    """Camera rays of random (pixel, sample), random rays through the scene box, and cosine bounce rays from the
    oracle's camera-ray hit points (8 floats per ray)."""
    p = setup.params()
    rng = np.random.default_rng(seed)
    npix = p.w * p.h
    cam = []
    for k in range(n // 2):
        o, d = po.camera_ray(p, int(rng.integers(npix)), int(rng.integers(256)))
        cam.append(np.concatenate([o, [0.0], d, [1e30]]))
    cam = np.array(cam, np.float32)
    v = setup.array(cwp.ARR_VERTICES)
    lo, hi = v.min(0), v.max(0)
    m = n // 4
    o = rng.uniform(lo - 0.1, hi + 0.1, (m, 3)).astype(np.float32)
    d = rng.normal(size=(m, 3)).astype(np.float32); d /= np.linalg.norm(d, axis=1, keepdims=True)
    rnd = np.concatenate([o, np.zeros((m, 1)), d, np.full((m, 1), 1e30)], 1).astype(np.float32)
    return cam, rnd


def bounce_rays(sc, cam, p, m, seed):
    h = sc.trace(cam)
    hit = h[:, 3].view(np.int32) >= 0
    rng = np.random.default_rng(seed)
    pts = cam[hit, :3] + cam[hit, 4:7] * h[hit, :1]
    pts = pts[:m]
    d = rng.normal(size=pts.shape).astype(np.float32); d /= np.linalg.norm(d, axis=1, keepdims=True)
    d[:, 2] = np.abs(d[:, 2])
    o = pts + d * np.float32(p.eps) + np.array([0, 0, p.eps], np.float32)
    return np.concatenate([o, np.zeros((len(o), 1)), d, np.full((len(o), 1), 1e30)], 1).astype(np.float32)


SCENES = {
    "boxes": ["--gen", "boxes", "--width", "96", "--height", "96"],
    "terrain": ["--gen", "terrain", "--width", "128", "--height", "72"],
}


@pytest.mark.parametrize("scene", ["boxes", "terrain"])
@pytest.mark.parametrize("variant", [8, 4, 2])
@pytest.mark.parametrize("tri", ["woop", "watertight", "mt"])
def test_trace_matches_oracle(scene, variant, tri):
    print("case", scene, variant, tri, flush=True)
    s = cwp.Setup(SCENES[scene] + ["--variant", "bvh%d" % variant, "--tri", tri, "--batch", "1"])
    r = cwp.Renderer(s)
    sc, _ = po.scene_from_setup(s)
    cam, rnd = scene_rays(s, 60000, 1)
    bnc = bounce_rays(sc, cam, s.params(), 30000, 2)
    rays = np.concatenate([cam, rnd, bnc])
    g, o = r.trace(rays), sc.trace(rays)
    gi, oi = g[:, 3].view(np.int32), o[:, 3].view(np.int32)
    agree = (gi == oi).mean()
    assert (oi >= 0).sum() > 0.2 * len(rays)
    assert agree >= 0.9999, agree
    both = (gi == oi) & (oi >= 0)
    rel = np.abs(g[both, 0] - o[both, 0]) / np.maximum(np.abs(o[both, 0]), 1e-30)
    assert rel.max() <= 1e-5
    ga, oa = r.trace(rays, any_hit=True), sc.trace(rays, any_hit=True)
    assert ((ga[:, 3].view(np.int32) >= 0) == (oa[:, 3].view(np.int32) >= 0)).mean() >= 0.9999
    # any hit reports occlusion exactly when a closest hit exists
    assert ((ga[:, 3].view(np.int32) >= 0) == (gi >= 0)).mean() >= 0.9999


def test_trace_empty_and_single():
    s = cwp.Setup(SCENES["boxes"] + ["--batch", "1"])
    r = cwp.Renderer(s)
    r.trace_ptr(0, 0, False, 0)   # n = 0 is a no-op
    one = np.array([[0.5, -3.0, 0.3, 0.0, 0.0, 1.0, 0.0, 1e30]], np.float32)
    g = r.trace(one)
    sc, _ = po.scene_from_setup(s)
    np.testing.assert_array_equal(g.view(np.int32)[:, 3], sc.trace(one).view(np.int32)[:, 3])


@pytest.mark.parametrize("scene,tri,variant", [("boxes", "woop", 8), ("boxes", "watertight", 2), ("terrain", "woop", 8),
                                               ("terrain", "watertight", 8), ("terrain", "woop", 4)])
def test_render_matches_oracle(scene, tri, variant):
    print("case", scene, tri, variant, flush=True)
    spp, batch, bounces = 16, 8, 4
    s = cwp.Setup(SCENES[scene] + ["--variant", "bvh%d" % variant, "--tri", tri, "--batch", str(batch),
                                   "--bounces", str(bounces), "--spp", str(spp)])
    r = cwp.Renderer(s)
    st = r.render(spp)
    img = r.accum()[..., :3] / spp
    sc, _ = po.scene_from_setup(s)
    acc, ost = sc.render(s.params(), spp, batch=batch, bounces=bounces)
    ref = acc[..., :3] / spp
    assert psnr(img, ref) >= 50.0
    assert st["rays_closest"] == pytest.approx(ost[0], rel=1e-3)
    assert st["rays_shadow"] == pytest.approx(ost[1], rel=1e-3)
    assert st["rays_miss"] == pytest.approx(ost[2], rel=1e-3)
    assert st["shadow_unoccluded"] == pytest.approx(ost[3], rel=1e-3)
    # most pixels are bit-identical: the path arithmetic is the oracle's, only libm sin/cos may differ in an ulp
    assert np.mean(np.all(img == ref, axis=-1)) > 0.5


def test_thin_lens_render_matches_oracle():
    """Thin-lens camera (concentric-disk lens sample, focus plane; reference: src/render/wavefront.cu:43-57) against the
    oracle's camera at >= 50 dB; the lens visibly changes the image against the pinhole one."""
    spp, batch, bounces = 16, 16, 3
    base = SCENES["terrain"] + ["--batch", str(batch), "--bounces", str(bounces), "--spp", str(spp)]
    s = cwp.Setup(base + ["--aperture", "0.02", "--focus_dist", "0.8"])
    p = s.params()
    assert p.lens_r == pytest.approx(0.02) and p.focus_d == pytest.approx(0.8)
    r = cwp.Renderer(s)
    st = r.render(spp)
    img = r.accum()[..., :3] / spp
    r.close()
    sc, _ = po.scene_from_setup(s)
    acc, ost = sc.render(p, spp, batch=batch, bounces=bounces)
    assert psnr(img, acc[..., :3] / spp) >= 50.0
    assert st["rays_closest"] == pytest.approx(ost[0], rel=1e-3)
    pin = cwp.Renderer(cwp.Setup(base))
    pin.render(spp)
    pimg = pin.accum()[..., :3] / spp
    pin.close()
    assert psnr(img, pimg) < 40.0


@pytest.mark.parametrize("scene,variant", [("terrain", 8), ("boxes", 2), ("boxes", 4)])
def test_kernel_schedules_give_the_same_image(scene, variant):
    """Fused trace+shade, split trace/shade, split with persistent refilling lanes, and the shadow kernel with and
    without refilling compute every path with the same arithmetic in the same order per path: images and counters are
    bit-identical, and match the oracle."""
    print("case", scene, variant, flush=True)
    spp, batch, bounces = 8, 8, 5
    base = SCENES[scene] + ["--variant", "bvh%d" % variant, "--batch", str(batch), "--bounces", str(bounces),
                            "--spp", str(spp)]
    out = {}
    schedules = [("fused", ["--fuse", "1", "--shadow_refill", "0"]), ("fused-refill", ["--fuse", "1"]),
                 ("split", ["--fuse", "0", "--shadow_refill", "0"]), ("persistent", ["--persistent", "1"]),
                 ("fused-persistent", ["--persistent", "2"])]
    for name, extra in schedules:
        s = cwp.Setup(base + extra)
        r = cwp.Renderer(s)
        st = r.render(spp)
        out[name] = (r.accum(), st)
        assert r.frame_json(st)["fuse"] == (1 if name.startswith("fused") else 0)
        r.close()
    img0, st0 = out["fused"]
    for name in ("fused-refill", "split", "persistent", "fused-persistent"):
        img, st = out[name]
        np.testing.assert_array_equal(img, img0, err_msg=name)
        for k in ("rays_closest", "rays_shadow", "rays_miss", "rays_cont", "shadow_unoccluded", "camera_miss"):
            assert st[k] == st0[k], (name, k)
    assert out["split"][1]["t_shade_ms"] > 0 and out["fused"][1]["t_shade_ms"] == 0
    sc, _ = po.scene_from_setup(cwp.Setup(base))
    acc, _ = sc.render(cwp.Setup(base).params(), spp, batch=batch, bounces=bounces)
    assert psnr(img0[..., :3] / spp, acc[..., :3] / spp) >= 50.0


SORT_CASES = [
    ("terrain-bvh8", SCENES["terrain"] + ["--variant", "bvh8"]),
    ("boxes-bvh2", SCENES["boxes"] + ["--variant", "bvh2"]),
    ("terrain-split", SCENES["terrain"] + ["--fuse", "0"]),
    ("terrain-refill", SCENES["terrain"] + ["--persistent", "2"]),
    ("boxes-full", SCENES["boxes"] + ["--gen_materials", "3", "--gen_env", "64,32,2,1", "--graph", "0"]),
    ("instances", ["--gen", "instances", "--gen_small", "1", "--gen_instances", "40", "--width", "96", "--height", "64"]),
    ("terrain-graph", SCENES["terrain"] + ["--graph", "1"]),
    ("terrain-256spp", SCENES["terrain"] + ["--variant", "bvh8", "--width", "48", "--height", "32"]),
]


@pytest.mark.parametrize("name,args", SORT_CASES, ids=[c[0] for c in SORT_CASES])
def test_sorted_continuations_give_the_same_image(name, args):
    """--sort 1 regroups each bounce's continuations by direction bucket within each partition, --sort 2 the first
    bounce's by elevation bucket within each pixel's samples (fused kernel only; single-level scenes group them inside
    the depth-0 kernel, two-level scenes in a separate pass), before they are traced. Every path
    keeps its own state and radiance, so the image and every counter are bit-identical to the unsorted schedule (a
    lost, doubled or misplaced path id would change both), on the first frame and on a second frame that reuses the
    counters."""
    print("case", name, flush=True)
    spp, batch = (256, 256) if name == "terrain-256spp" else (16, 8)   # 256: a sort tile is one pixel's samples
    base = args + ["--batch", str(batch), "--bounces", "6", "--spp", str(spp)]
    out = {}
    fused = "--fuse" not in args and "--persistent" not in args
    for sort in (0, 1, 2) if fused else (0, 1):
        r = cwp.Renderer(cwp.Setup(base + ["--sort", str(sort)]))
        frames = []
        for f in range(2):
            if name == "instances":
                r.set_frame(f)
            st = r.render(spp, sample_base=f * spp)
            frames.append((r.accum(), st))
        out[sort] = frames
        j = r.frame_json(frames[-1][1])
        assert j["sort"] == sort
        timed = "--graph" not in args or "0" == args[args.index("--graph") + 1]
        separate = sort == 1 or (sort == 2 and name == "instances")   # single-level scenes fuse sort 2 into depth 0
        assert (j["t_sort_ms"] > 0) == (separate and timed)
        r.close()
    for sort in sorted(out)[1:]:
        for f in range(2):
            (a, sa), (b, sb) = out[0][f], out[sort][f]
            np.testing.assert_array_equal(b, a, err_msg="sort %d frame %d" % (sort, f))
            for k in ("rays_closest", "rays_shadow", "rays_miss", "rays_cont", "shadow_unoccluded", "camera_miss"):
                assert sa[k] == sb[k], (sort, f, k)
    assert out[0][0][1]["rays_cont"] > 0


def test_render_progressive_and_deterministic():
    s = cwp.Setup(SCENES["boxes"] + ["--batch", "8", "--bounces", "3"])
    r = cwp.Renderer(s)
    r.render(16)
    one = r.accum()
    r.render(8, 0, True)
    r.render(8, 8, False)
    two = r.accum()
    np.testing.assert_array_equal(one, two)   # N calls of K samples == one call of N*K samples, bit for bit
    r.render(16)
    np.testing.assert_array_equal(one, r.accum())


def test_batch_size_does_not_change_the_image():
    out = []
    for batch in (4, 16):
        s = cwp.Setup(SCENES["terrain"] + ["--batch", str(batch), "--bounces", "3"])
        r = cwp.Renderer(s)
        r.render(16)
        out.append(r.accum())
    # per-pass sums differ only in float summation grouping
    assert psnr(out[0][..., :3] / 16, out[1][..., :3] / 16) > 80


def test_plane_closed_form_on_gpu(tmp_path):
    obj = tmp_path / "plane.obj"
    obj.write_text("v -1000 -1000 0\nv 1000 -1000 0\nv 1000 1000 0\nv -1000 1000 0\nf 1 2 3\nf 1 3 4\n")
    a, sky = 0.6, (0.4, 0.5, 0.6)
    s = cwp.Setup(["--mesh", str(obj), "--c2w", "1,0,0,0,0,0.7071068,-0.7071068,-2,0,0.7071068,0.7071068,1,0,0,0,1",
                   "--width", "48", "--height", "48", "--albedo", str(a), str(a), str(a), "--sky", *map(str, sky),
                   "--sun_rgb", "0", "0", "0", "--filter", "box", "--batch", "8", "--spp", "8"])
    r = cwp.Renderer(s)
    st = r.render(8)
    img = r.accum()[..., :3] / 8
    f = (1.0 - img / np.array(sky, np.float32)) / (1.0 - a)
    np.testing.assert_allclose(f * 8, np.round(f * 8), atol=1e-4)
    assert np.isclose(f[..., 0], 1.0, atol=1e-5).sum() > 0.3 * 48 * 48
    assert st["rays_closest"] == 8 * 48 * 48 + np.round(f[..., 0] * 8).sum()


def test_frame_json_has_reference_keys():
    s = cwp.Setup(SCENES["boxes"] + ["--batch", "4", "--spp", "4"])
    r = cwp.Renderer(s)
    r.render(4)
    j = r.frame_json()
    for k in ("rays_closest", "rays_shadow", "rays_total", "t_trace_ms", "t_frame_ms", "mrays_s_trace", "mrays_s_e2e",
              "avg_path_len", "vram_MB", "bvh_nodes", "bvh_bytes", "bvh_sah", "variant", "tri", "batch", "spp"):
        assert k in j
    assert j["rays_total"] == j["rays_closest"] + j["rays_shadow"]


def test_renderer_from_caller_buffers_matches_setup_renderer():
    """cwp_create_scene over the setup's own arrays renders the same image bit for bit as cwp_create."""
    s = cwp.Setup(SCENES["boxes"] + ["--batch", "8", "--bounces", "4", "--spp", "8"])
    a = cwp.Renderer(s)
    a.render(8)
    ref = a.accum()
    a.close()
    opts = cwp.Options.default(batch=8, bounces=4)
    r = cwp.Renderer.from_scene(8, s.array(cwp.ARR_NODES), s.array(cwp.ARR_CORNERS), s.array(cwp.ARR_ALBEDO),
                                s.params(), opts, woop=s.array(cwp.ARR_WOOP))
    st = r.render(8)
    np.testing.assert_array_equal(r.accum(), ref)
    assert st["rays_closest"] > 0
    with pytest.raises(cwp.CwpError, match="frame_json|null"):
        r.frame_json(st)
    r.close()


@pytest.mark.parametrize("full", [0, 1])
def test_graph_frames_equal_stream_frames(full):
    """frames replayed from a captured graph equal frames launched one kernel at a time, bit for bit, on every
    replay, with other work (a different frame, a parameter change) between the replays"""
    args = SCENES["terrain"] + ["--batch", "8", "--bounces", "4", "--spp", "16"]
    if full:
        args += ["--gen_materials", "3", "--gen_env", "64,32,2,1"]
    ref = cwp.Renderer(cwp.Setup(args + ["--graph", "0"]))
    ref.render(16)
    a = ref.accum()
    ref.render(8, sample_base=16)
    b = ref.accum()
    g = cwp.Renderer(cwp.Setup(args + ["--graph", "1"]))
    for rep in range(3):
        st = g.render(16)
        np.testing.assert_array_equal(g.accum(), a, err_msg="replay %d" % rep)
        assert g.frame_json(st)["graph"] == 1 and st["t_bounce_ms"] == 0
        g.render(8, sample_base=16)            # another captured frame in between
        np.testing.assert_array_equal(g.accum(), b)
    p = g.setup.params()
    g.set_params(p)                            # parameters are read through a pointer: the graphs stay valid
    g.render(16)
    np.testing.assert_array_equal(g.accum(), a)


TAIL_CASES = [
    ("terrain", SCENES["terrain"] + ["--variant", "bvh8"], (0, 2)),
    ("terrain-watertight", SCENES["terrain"] + ["--variant", "bvh8", "--tri", "watertight"], (0,)),
    ("terrain-mt-stack16", SCENES["terrain"] + ["--variant", "bvh8", "--tri", "mt", "--stack", "16"], (0,)),
    ("boxes-full", SCENES["boxes"] + ["--gen_materials", "3", "--gen_env", "64,32,2,1", "--graph", "0"], (0, 2)),
    ("terrain-graph", SCENES["terrain"] + ["--graph", "1"], (0, 1, 2)),
    ("terrain-256spp", SCENES["terrain"] + ["--width", "48", "--height", "32"], (2,)),
    ("instances", ["--gen", "instances", "--gen_small", "1", "--gen_instances", "40", "--width", "96", "--height", "64"],
     (0, 2)),
    ("instances-full", ["--gen", "instances", "--gen_small", "1", "--gen_instances", "30", "--width", "64", "--height", "48",
                        "--gen_materials", "2", "--gen_env", "64,32,1,3", "--graph", "0"], (0,)),
]


@pytest.mark.parametrize("name,args,sorts", TAIL_CASES, ids=[c[0] for c in TAIL_CASES])
def test_straggler_hand_off_gives_the_same_image(name, args, sorts):
    """--tail N: a trace+shade or shadow wave whose tracing lanes drop below N hands them to a tail kernel, which resumes
    each traversal from its record (hit so far, node group, stack) and, with --tail_levels 2, hands its own stragglers
    on to a second tail kernel. The traversal is the same, so images and every counter are bit-identical to the plain
    kernels, at N = 64 (nearly every ray handed off once a lane finishes, and on again) and N = 8, with one and two
    levels, on two frames in a row (the tail counters are cleared per pass)."""
    spp, batch = (256, 256) if name == "terrain-256spp" else (16, 8)
    base = args + ["--batch", str(batch), "--bounces", "6", "--spp", str(spp)]
    for sort in sorts:
        out = {}
        for tail, levels in ((0, 1), (8, 2), (64, 2), (8, 1), (64, 1)):
            print("case", name, "sort", sort, "tail", tail, "levels", levels, flush=True)
            r = cwp.Renderer(cwp.Setup(base + ["--sort", str(sort), "--tail", str(tail), "--tail_levels", str(levels)]))
            frames = []
            for f in range(2):
                if name.startswith("instances"):
                    r.set_frame(f)
                st = r.render(spp, sample_base=f * spp)
                frames.append((r.accum(), st))
            out[tail, levels] = frames
            r.close()
        for key in out:
            for f in range(2):
                (a, sa), (b, sb) = out[0, 1][f], out[key][f]
                np.testing.assert_array_equal(a, b, err_msg="tail %s frame %d" % (key, f))
                for k in ("rays_closest", "rays_shadow", "rays_miss", "rays_cont", "shadow_unoccluded", "camera_miss"):
                    assert sa[k] == sb[k], (key, f, k)


def test_straggler_hand_off_rejects_unsupported_schedules():
    for extra in (["--variant", "bvh2"], ["--fuse", "0"], ["--persistent", "2"]):
        s = cwp.Setup(SCENES["boxes"] + ["--tail", "8"] + extra)
        with pytest.raises(cwp.CwpError):
            cwp.Renderer(s)
    s = cwp.Setup(["--gen", "instances", "--gen_small", "1", "--gen_instances", "10", "--width", "32", "--height", "32",
                   "--tlas", "bvh2", "--tail", "8"])
    with pytest.raises(cwp.CwpError):   # the hand-off resumes BVH8 TLAS steps
        cwp.Renderer(s)
