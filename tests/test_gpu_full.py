"""GPU parity of the full material integrator (OpenPBR-subset materials, emissive triangles sampled by power, the
environment map sampled by luminance, MIS, Russian roulette) against the oracle, through the C ABI.

Bars: images >= 50 dB PSNR against the oracle (sRGB, clip [0, 1]) at the same samples; ray counters within 1 %
(a path may take another branch where a device transcendental differs from the host's in the last place); the
closed forms the oracle is pinned on (tests/test_oracle_full.py) hold on the GPU too.
"""
import math

import numpy as np
import pytest

import cwp
import pyoracle as po

pytestmark = pytest.mark.gpu

TERRAIN = ["--gen", "terrain", "--gen_grid", "60", "--gen_boxes", "24", "--gen_materials", "5", "--width", "64",
           "--height", "48", "--filter", "box"]
BOXES = ["--gen", "boxes", "--gen_boxes", "96", "--gen_materials", "7", "--width", "56", "--height", "56"]
ENV = ["--gen_env", "128,64,2,3"]


def psnr(a, b):
    def srgb(x):
        x = np.clip(x, 0, 1)
        return np.where(x <= 0.0031308, 12.92 * x, 1.055 * np.power(x, 1 / 2.4) - 0.055)
    mse = float(np.mean((srgb(a.astype(np.float64)) - srgb(b.astype(np.float64))) ** 2))
    return 99.0 if mse == 0 else 10 * math.log10(1.0 / mse)


CASES = [
    ("terrain-materials-bvh8", TERRAIN + ["--variant", "bvh8"]),
    ("terrain-env-mis-bvh8", TERRAIN + ENV + ["--variant", "bvh8"]),
    ("terrain-env-nee", TERRAIN + ENV + ["--variant", "bvh8", "--light_sampling", "nee"]),
    ("terrain-env-bsdf", TERRAIN + ENV + ["--variant", "bvh8", "--light_sampling", "bsdf"]),
    ("terrain-env-rr-bvh2-watertight", TERRAIN + ENV + ["--variant", "bvh2", "--tri", "watertight", "--rr", "1", "--rr_depth", "2"]),
    ("boxes-uniform-env-bvh4", BOXES + ["--gen_env", "64,32,1,2", "--env_sampling", "uniform", "--variant", "bvh4", "--tri", "mt"]),
    ("terrain-plain-mix-bh-filter", TERRAIN + ENV + ["--variant", "bvh8", "--gd_energy", "0", "--filter", "bh"]),
]


@pytest.mark.parametrize("name,args", CASES, ids=[c[0] for c in CASES])
def test_full_render_matches_oracle(name, args):
    print("case", name, flush=True)
    spp, bounces = 32, 4
    s = cwp.Setup(args + ["--bounces", str(bounces), "--spp", str(spp), "--batch", str(spp)])
    assert s.info["full"] == 1
    r = cwp.Renderer(s)
    st = r.render(spp)
    img = r.accum()[..., :3] / spp
    assert r.frame_json(st)["integrator"] == "full"
    r.close()
    sc, _ = po.scene_from_setup(s)
    acc, ost = sc.render(s.params(), spp, batch=spp, bounces=bounces)
    ref = acc[..., :3] / spp
    assert np.isfinite(img).all()
    assert psnr(img, ref) >= 50.0, (name, psnr(img, ref))
    assert st["rays_closest"] == pytest.approx(ost[0], rel=1e-2)
    assert st["rays_shadow"] == pytest.approx(ost[1], rel=1e-2)
    assert st["shadow_unoccluded"] == pytest.approx(ost[3], rel=1e-2)


def test_full_persistent_kernel_gives_the_same_image():
    """the fused kernel with refilled lanes computes every path with the same arithmetic: the same image, counters"""
    base = TERRAIN + ENV + ["--bounces", "4", "--spp", "8", "--batch", "8"]
    out = []
    for extra in ([], ["--persistent", "2"]):
        r = cwp.Renderer(cwp.Setup(base + extra))
        st = r.render(8)
        out.append((r.accum(), st))
        r.close()
    np.testing.assert_array_equal(out[0][0], out[1][0])
    for k in ("rays_closest", "rays_shadow", "rays_miss", "shadow_unoccluded"):
        assert out[0][1][k] == out[1][1][k], k


def test_full_kernel_is_deterministic_and_progressive():
    """the same image twice, and 2 x 8 samples accumulated equal 16 samples in two passes of 8"""
    s = cwp.Setup(TERRAIN + ENV + ["--bounces", "4", "--spp", "16", "--batch", "8"])
    r = cwp.Renderer(s)
    r.render(16)
    a = r.accum()
    r.render(16)
    np.testing.assert_array_equal(r.accum(), a)
    r.render(8, sample_base=0)
    r.render(8, sample_base=8, clear=False)
    np.testing.assert_array_equal(r.accum(), a)


def test_full_plane_closed_form_on_gpu(tmp_path):
    """a Lambert plane of albedo a under a constant sky L, no sun: a sample that hits the plane is a * L (to float
    rounding), one that misses is L, so each pixel is L (1 - f (1 - a)) with f = hits / 8"""
    obj = tmp_path / "plane.obj"
    obj.write_text("v -1000 -1000 0\nv 1000 -1000 0\nv 1000 1000 0\nv -1000 1000 0\nf 1 2 3\nf 1 3 4\n")
    a, sky = 0.6, (0.4, 0.5, 0.6)
    s = cwp.Setup(["--mesh", str(obj), "--c2w", "1,0,0,0,0,0.7071068,-0.7071068,-2,0,0.7071068,0.7071068,1,0,0,0,1",
                   "--width", "48", "--height", "48", "--albedo", str(a), str(a), str(a), "--sky", *map(str, sky),
                   "--sun_rgb", "0", "0", "0", "--filter", "box", "--batch", "8", "--spp", "8", "--integrator", "full",
                   "--bounces", "0"])
    r = cwp.Renderer(s)
    r.render(8)
    img = r.accum()[..., :3] / 8
    f = (1.0 - img / np.array(sky, np.float32)) / (1.0 - a)
    np.testing.assert_allclose(f * 8, np.round(f * 8), atol=2e-3)
    assert np.isclose(f[..., 0], 1.0, atol=1e-4).sum() > 0.3 * 48 * 48


def _full_fields(s):
    i = s.info
    return {"mats": s.array(cwp.ARR_MATERIALS), "tri_light_pdf": s.array(cwp.ARR_TRI_LIGHT_PDF),
            "em_tri": s.array(cwp.ARR_EM_TRI), "em_cdf": s.array(cwp.ARR_EM_CDF),
            "env_texels": s.array(cwp.ARR_ENV_TEX).reshape(i["env_h"], i["env_w"], 4) if i["env_w"] else None,
            "env_rot_deg": i["env_rot_deg"], "sun_on": i["sun_on"], "light_mode": i["light_mode"], "rr": i["rr"],
            "rr_depth": i["rr_depth"], "gd_energy": i["gd_energy"], "env_uniform": i["env_uniform"]}


def test_full_scene_from_caller_buffers_matches_setup_renderer():
    """cwp_create_scene with the full integrator's fields (materials, emitters, environment) over the setup's own
    arrays renders the same image bit for bit as cwp_create."""
    args = TERRAIN + ["--gen_materials", "3", "--gen_env", "64,32,2,1", "--batch", "8", "--bounces", "4", "--spp", "8",
                      "--graph", "0"]
    s = cwp.Setup(args)
    assert s.info["full"] and s.info["num_emitters"] > 0
    a = cwp.Renderer(s)
    a.render(8)
    ref = a.accum()
    a.close()
    opts = cwp.Options.default(batch=8, bounces=4, graph=0)
    r = cwp.Renderer.from_scene(8, s.array(cwp.ARR_NODES), s.array(cwp.ARR_CORNERS), s.array(cwp.ARR_ALBEDO), s.params(),
                                opts, woop=s.array(cwp.ARR_WOOP), full=_full_fields(s))
    r.render(8)
    np.testing.assert_array_equal(r.accum(), ref)
    r.close()


def test_set_materials_and_environment_equal_fresh_scenes():
    """Runtime updates (reference: WavefrontRenderer::setMaterials / setEnvironment): a renderer whose environment is
    replaced renders exactly what a renderer created with it renders; new materials change the image; changing an
    emission is rejected; the legacy integrator takes the base colours as albedo (equal to a scene created with them)."""
    base = TERRAIN + ["--batch", "8", "--bounces", "4", "--spp", "8", "--graph", "0"]
    sa = cwp.Setup(base + ["--gen_env", "64,32,2,1"])
    sb = cwp.Setup(base + ["--gen_env", "64,32,2,9", "--env_rot", "30"])
    ra, rb = cwp.Renderer(sa), cwp.Renderer(sb)
    rb.render(8)
    want = rb.accum()
    rb.close()
    ib = sb.info
    assert ib["env_rot_deg"] == pytest.approx(30.0)
    ra.set_environment(sb.array(cwp.ARR_ENV_TEX).reshape(ib["env_h"], ib["env_w"], 4), rot_deg=ib["env_rot_deg"], sun_on=True)
    ra.render(8)
    np.testing.assert_array_equal(ra.accum(), want)
    m = sa.array(cwp.ARR_MATERIALS).copy()
    m[:, 7] = np.clip(m[:, 7] * 0.5, 0, 1)   # roughness
    m[:, 0:3] *= 0.5                         # base colour
    ra.set_materials(m)
    ra.render(8)
    assert not np.array_equal(ra.accum(), want)
    bad = m.copy()
    bad[:, 8] += 1.0                         # emission
    with pytest.raises(cwp.CwpError, match="emission"):
        ra.set_materials(bad)
    with pytest.raises(cwp.CwpError, match="texels"):
        ra.set_environment(np.zeros((4, 4, 4), np.float32))
    ra.close()
    legacy = ["--gen", "boxes", "--gen_boxes", "40", "--width", "40", "--height", "32", "--batch", "4", "--bounces", "3",
              "--spp", "4"]
    lg = cwp.Setup(legacy)
    rl = cwp.Renderer(lg)
    mats = np.zeros((lg.info["num_materials"], 16), np.float32)
    mats[:, :3] = (0.3, 0.4, 0.5)
    mats[:, 7] = 1.0
    mats[:, 11] = 1.5                        # ior
    rl.set_materials(mats)
    rl.render(4)
    got = rl.accum()
    rl.close()
    r2 = cwp.Renderer(cwp.Setup(legacy + ["--albedo", "0.3", "0.4", "0.5"]))
    r2.render(4)
    np.testing.assert_array_equal(got, r2.accum())
    r2.close()
    with pytest.raises(cwp.CwpError):
        cwp.Renderer(lg).set_environment(None, 10.0)   # legacy integrator: no environment
