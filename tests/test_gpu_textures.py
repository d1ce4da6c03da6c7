"""GPU parity of textured shading (base colour, emission, normal, metalness and roughness maps, bilinear / nearest
lookups, terminator shadowing) against the oracle, through the C ABI: generated textures on the terrain with
materials and an environment, an OBJ / MTL model with PNG textures, a two-level scene of textured meshes against its
flattened form, a scene built from caller buffers, and textured scenes in the job server.

Bars: images >= 50 dB PSNR against the oracle (sRGB, clip [0, 1]) at the same samples, ray counters within 1 %;
bit-identical images where only the route differs (caller buffers vs setup, a job alone vs in a mix).
"""
import json
import math

import numpy as np
import pytest

import cwp
import pyoracle as po
from test_textures import write_png

pytestmark = pytest.mark.gpu

TERRAIN = ["--gen", "terrain", "--gen_grid", "60", "--gen_boxes", "24", "--gen_materials", "5", "--width", "64",
           "--height", "48", "--filter", "box", "--gen_textures", "128"]
ENV = ["--gen_env", "128,64,2,3"]


def psnr(a, b):
    def srgb(x):
        x = np.clip(x, 0, 1)
        return np.where(x <= 0.0031308, 12.92 * x, 1.055 * np.power(x, 1 / 2.4) - 0.055)
    mse = float(np.mean((srgb(a.astype(np.float64)) - srgb(b.astype(np.float64))) ** 2))
    return 99.0 if mse == 0 else 10 * math.log10(1.0 / mse)


def render_both(args, spp=32, bounces=4):
    s = cwp.Setup(args + ["--bounces", str(bounces), "--spp", str(spp), "--batch", str(spp)])
    assert s.info["textured"] == 1 and s.info["full"] == 1
    r = cwp.Renderer(s)
    st = r.render(spp)
    img = r.accum()[..., :3] / spp
    r.close()
    sc, _ = po.scene_from_setup(s)
    acc, ost = sc.render(s.params(), spp, batch=spp, bounces=bounces)
    return s, img, st, acc[..., :3] / spp, ost


CASES = [
    ("terrain-linear", TERRAIN + ENV),
    ("terrain-nearest", TERRAIN + ENV + ["--tex_filter", "nearest"]),
    ("terrain-bump-shadow", TERRAIN + ENV + ["--bump_shadow", "1"]),
    ("terrain-base-roughness-only", TERRAIN + ENV + ["--textures", "base,roughness"]),
    ("terrain-sun-nee-bvh2", TERRAIN + ["--variant", "bvh2", "--light_sampling", "nee"]),
]


@pytest.mark.parametrize("name,args", CASES, ids=[c[0] for c in CASES])
def test_textured_render_matches_oracle(name, args):
    print("case", name, flush=True)
    s, img, st, ref, ost = render_both(args)
    assert np.isfinite(img).all()
    assert psnr(img, ref) >= 50.0, (name, psnr(img, ref))
    assert st["rays_closest"] == pytest.approx(ost[0], rel=1e-2)
    assert st["rays_shadow"] == pytest.approx(ost[1], rel=1e-2)


def test_textures_change_the_image():
    """the maps are used: without them (--textures none of the kinds) the image differs clearly"""
    a = render_both(TERRAIN + ENV, spp=8)[1]
    s = cwp.Setup(TERRAIN + ENV + ["--textures", "emission", "--bounces", "4", "--spp", "8", "--batch", "8"])
    assert s.info["textured"] == 0   # no emission maps bound: no textures
    r = cwp.Renderer(s)
    r.render(8)
    b = r.accum()[..., :3] / 8
    r.close()
    assert psnr(a, b) < 30.0


def heightfield_obj(tmp_path, n=24, name="hf"):
    """a wavy n x n quad patch over [-1, 1]^2 with per-vertex normals and uvs (2 repeats), its MTL binding a PNG base
    colour map (sRGB), a PNG normal map, a PNG roughness map (G) and an emission map on a second material"""
    rng = np.random.default_rng(4)
    base = rng.integers(30, 250, (16, 16, 3))
    y, x = np.mgrid[0:32, 0:32]
    nx, ny = 0.4 * np.cos(x / 32 * 2 * np.pi * 3), 0.4 * np.sin(y / 32 * 2 * np.pi * 2)
    nz = np.sqrt(1 - nx * nx - ny * ny)
    nrm = np.stack([(nx * 0.5 + 0.5) * 255, (ny * 0.5 + 0.5) * 255, (nz * 0.5 + 0.5) * 255], -1).round().astype(int)
    rough = np.stack([np.zeros((8, 8)), rng.integers(60, 255, (8, 8)), np.zeros((8, 8))], -1).astype(int)
    emis = rng.integers(0, 255, (4, 4, 3))
    write_png(tmp_path / ("%s_base.png" % name), base, 2, 8)
    write_png(tmp_path / ("%s_n.png" % name), nrm, 2, 8, interlace=True)
    write_png(tmp_path / ("%s_r.png" % name), rough, 2, 8)
    write_png(tmp_path / ("%s_e.png" % name), emis, 2, 8, level=0)
    (tmp_path / ("%s.mtl" % name)).write_text(
        "newmtl surf\nKd 1 1 1\nmap_Kd %s_base.png\nmap_Bump -bm 0.8 %s_n.png\nmap_Pr -imfchan g %s_r.png\n"
        "newmtl glow\nKd 0.5 0.5 0.5\nKe 4 3 2\nmap_Ke %s_e.png\n" % (name, name, name, name))
    lines = ["mtllib %s.mtl" % name]
    for j in range(n + 1):
        for i in range(n + 1):
            u, v = i / n, j / n
            px, py = 2 * u - 1, 2 * v - 1
            z = 0.15 * math.sin(3 * px) * math.cos(2 * py)
            dzdx, dzdy = 0.45 * math.cos(3 * px) * math.cos(2 * py), -0.3 * math.sin(3 * px) * math.sin(2 * py)
            lines.append("v %.6f %.6f %.6f" % (px, py, z))
            lines.append("vn %.6f %.6f 1" % (-dzdx, -dzdy))
            lines.append("vt %.6f %.6f" % (2 * u, 2 * v))
    lines.append("usemtl surf")
    for j in range(n):
        for i in range(n):
            a = j * (n + 1) + i + 1
            b, c, d = a + 1, a + n + 2, a + n + 1
            if i == n - 1 and j == n - 1:
                lines.append("usemtl glow")
            lines.append("f %d/%d/%d %d/%d/%d %d/%d/%d" % (a, a, a, b, b, b, c, c, c))
            lines.append("f %d/%d/%d %d/%d/%d %d/%d/%d" % (a, a, a, c, c, c, d, d, d))
    (tmp_path / ("%s.obj" % name)).write_text("\n".join(lines) + "\n")
    return str(tmp_path / ("%s.obj" % name))


VIEW = ["--c2w", "1,0,0,0,0,0.6,-0.8,-2.6,0,0.8,0.6,2.4,0,0,0,1", "--hfov", "45", "--width", "56", "--height", "48"]


def test_obj_png_textures_match_oracle(tmp_path):
    obj = heightfield_obj(tmp_path)
    s, img, st, ref, ost = render_both(["--mesh", obj] + VIEW + ENV + ["--integrator", "full", "--bump_shadow", "1"])
    assert s.info["num_images"] == 4
    assert psnr(img, ref) >= 50.0, psnr(img, ref)
    assert st["rays_closest"] == pytest.approx(ost[0], rel=1e-2)


@pytest.mark.parametrize("integrator", ["legacy", "full"])
def test_smooth_normals_match_oracle(tmp_path, integrator):
    """per-vertex normals interpolated at the hit (smooth shading) in both integrators, without textures"""
    obj = heightfield_obj(tmp_path)
    s = cwp.Setup(["--mesh", obj] + VIEW + ["--integrator", integrator, "--textures", "none", "--bounces", "4",
                                            "--spp", "16", "--batch", "16"])
    assert s.params().has_normals == 1 and s.info["textured"] == 0
    r = cwp.Renderer(s)
    st = r.render(16)
    img = r.accum()[..., :3] / 16
    r.close()
    sc, _ = po.scene_from_setup(s)
    acc, ost = sc.render(s.params(), 16, batch=16, bounces=4)
    assert psnr(img, acc[..., :3] / 16) >= 50.0
    assert st["rays_closest"] == pytest.approx(ost[0], rel=1e-2)


def test_two_level_textured_matches_flattened(tmp_path):
    """textured meshes as instances (a rotated one, a mirrored one: tangents through the forward rows, handedness
    times the determinant's sign) against the same instances flattened into one mesh; the flattened scene against
    the oracle"""
    obj = heightfield_obj(tmp_path)
    scene = {"meshes": [{"file": obj}],
             "instances": [{"mesh": 0, "translate": [-1.1, 0, 0]},
                           {"mesh": 0, "translate": [1.1, 0, 0], "axis": [0, 0, 1], "angle_deg": 90},
                           {"mesh": 0, "translate": [0, 2.1, 0], "scale": [-1, 1, 1]}]}
    (tmp_path / "scene.json").write_text(json.dumps(scene))
    view = ["--c2w", "1,0,0,0,0,0.6,-0.8,-4.6,0,0.8,0.6,3.6,0,0,0,1", "--hfov", "60", "--width", "64", "--height", "48"]
    args = ["--scene", str(tmp_path / "scene.json")] + view + ENV + ["--integrator", "full", "--bounces", "3", "--spp", "16",
                                                                    "--batch", "16"]
    two = cwp.Setup(args)
    assert two.info["num_instances"] == 3 and two.info["textured"] == 1
    r = cwp.Renderer(two)
    r.render(16)
    a = r.accum()[..., :3] / 16
    r.close()
    flat = cwp.Setup(args + ["--flatten", "1"])
    assert flat.info["textured"] == 1 and flat.info["num_meshes"] == 0
    rf = cwp.Renderer(flat)
    rf.render(16)
    b = rf.accum()[..., :3] / 16
    rf.close()
    assert psnr(a, b) >= 45.0, psnr(a, b)
    sc, _ = po.scene_from_setup(flat)
    acc, _ = sc.render(flat.params(), 16, batch=16, bounces=3)
    assert psnr(b, acc[..., :3] / 16) >= 50.0


def test_textured_scene_from_caller_buffers_equals_setup():
    s = cwp.Setup(TERRAIN + ENV + ["--batch", "8", "--bounces", "3", "--spp", "8", "--graph", "0"])
    a = cwp.Renderer(s)
    a.render(8)
    want = a.accum()
    a.close()
    i = s.info
    full = {"mats": s.array(cwp.ARR_MATERIALS), "tri_light_pdf": s.array(cwp.ARR_TRI_LIGHT_PDF),
            "em_tri": s.array(cwp.ARR_EM_TRI), "em_cdf": s.array(cwp.ARR_EM_CDF),
            "env_texels": s.array(cwp.ARR_ENV_TEX).reshape(i["env_h"], i["env_w"], 4), "env_rot_deg": i["env_rot_deg"],
            "sun_on": i["sun_on"], "light_mode": i["light_mode"], "rr": i["rr"], "rr_depth": i["rr_depth"],
            "gd_energy": i["gd_energy"], "env_uniform": i["env_uniform"], "maps": s.array(cwp.ARR_MAPS),
            "uvs": s.array(cwp.ARR_UVS), "tangents": s.array(cwp.ARR_TANGENTS), "tex_images": s.array(cwp.ARR_TEX_IMAGES),
            "texels": s.array(cwp.ARR_TEXELS), "tex_nearest": i["tex_nearest"], "bump_shadow": i["bump_shadow"]}
    opts = cwp.Options.default(batch=8, bounces=3, graph=0)
    r = cwp.Renderer.from_scene(8, s.array(cwp.ARR_NODES), s.array(cwp.ARR_CORNERS), s.array(cwp.ARR_ALBEDO), s.params(),
                                opts, woop=s.array(cwp.ARR_WOOP), full=full)
    r.render(8)
    np.testing.assert_array_equal(r.accum(), want)
    r.close()
    bad = dict(full)
    bad["maps"] = full["maps"].copy()
    bad["maps"][0, 0] = 99   # an image that does not exist
    with pytest.raises(cwp.CwpError, match="image id"):
        cwp.Renderer.from_scene(8, s.array(cwp.ARR_NODES), s.array(cwp.ARR_CORNERS), s.array(cwp.ARR_ALBEDO), s.params(),
                                opts, woop=s.array(cwp.ARR_WOOP), full=bad)


def test_textured_jobs_on_the_server():
    """textures stay resident with their scene: a textured job alone equals the job in a mix, and equals the
    single-job renderer at the scene's resolution"""
    tex = cwp.Setup(TERRAIN + ENV + ["--bounces", "4", "--spp", "8", "--batch", "8", "--seed", "77"])
    plain = cwp.Setup(["--gen", "boxes", "--gen_boxes", "60", "--gen_materials", "3", "--width", "40", "--height", "40",
                       "--bounces", "4"])
    srv = cwp.Server(variant=8, bounces=4, max_paths=1 << 18, full=1)
    a, b = srv.add_scene(tex), srv.add_scene(plain)
    jobs = [(a, 64, 48, 8, 4, 77), (b, 40, 40, 8, 4, 5), (a, 32, 32, 4, 3, 9)]

    def run(js):
        srv.clear_jobs()
        ids = [srv.submit(*j[:5], seed=j[5]) for j in js]
        srv.run(batched=True)
        return [srv.image(i) for i in ids]
    mix = run(jobs)
    for k, j in enumerate(jobs):
        np.testing.assert_array_equal(run([j])[0], mix[k], err_msg="job %d" % k)
    r = cwp.Renderer(tex)
    r.render(8)
    np.testing.assert_array_equal(mix[0], r.accum())
    r.close()
    srv.close()
