"""Pins of the oracle's full material integrator (no GPU): BSDF sampling against its pdf (chi-square, as the
reference's tests/bsdf_tests.cpp), white-furnace albedos against a deterministic quadrature, Fresnel / F82 / GGX
closed forms, the environment and emitter light distributions, and closed-form scenes that exercise the MIS weights:

* a Lambert plane under a constant environment E: every plane pixel has expectation a * E;
* the same plane under a log-normal environment: a / pi * integral of L(w) cos(theta) over the upper hemisphere,
  computed by quadrature over the map's texels;
* a closed box whose walls all emit E with Lambert albedo a, bounce limit B: every pixel has expectation
  E (1 - a^(B+2)) / (1 - a); wrong MIS weights, light pdfs or emission handling bias it.

The product's host tables (emitter CDF, environment CDFs) are compared with the oracle's bit for bit.
"""
import math

import numpy as np
import pytest
from scipy import stats

import cwp
import pyoracle as po


def _dirs_grid(nc, nphi, sub):
    """cell-centred directions of an equal-solid-angle grid: cos(theta) in [-1, 1], phi in [0, 2 pi)"""
    c = -1.0 + (np.arange(nc * sub) + 0.5) * (2.0 / (nc * sub))
    ph = (np.arange(nphi * sub) + 0.5) * (2 * np.pi / (nphi * sub))
    C, PH = np.meshgrid(c, ph, indexing="ij")
    st = np.sqrt(np.maximum(0, 1 - C * C))
    return np.stack([st * np.cos(PH), st * np.sin(PH), C], -1).reshape(-1, 3).astype(np.float32)


def chi2_pvalue(mat, theta_o, eta, n=200_000, seed=7, nc=20, nphi=40, sub=24):
    wo = np.array([math.sin(math.radians(theta_o)), 0.0, math.cos(math.radians(theta_o))], np.float32)
    wi, w, pdf, ok = po.bsdf_sample_batch(mat, wo, eta, n, seed)
    good = (ok != 0) & (pdf > 0)
    obs = np.zeros(nc * nphi + 1)
    c = np.clip(wi[good, 2].astype(np.float64), -1, 1)
    ph = np.mod(np.arctan2(wi[good, 1], wi[good, 0]).astype(np.float64), 2 * np.pi)
    ic = np.minimum(nc - 1, ((c + 1) * 0.5 * nc).astype(int))
    ip = np.minimum(nphi - 1, (ph / (2 * np.pi) * nphi).astype(int))
    np.add.at(obs, ic * nphi + ip, 1)
    obs[-1] = (~good).sum()
    d = _dirs_grid(nc, nphi, sub)
    p = po.bsdf_pdf_batch(mat, wo, eta, d).astype(np.float64).reshape(nc, sub, nphi, sub)
    cell = (2.0 / nc / sub) * (2 * np.pi / nphi / sub)
    exp = np.zeros(nc * nphi + 1)
    exp[:-1] = p.sum(axis=(1, 3)).reshape(-1) * cell * n
    mass = exp[:-1].sum() / n
    exp[-1] = max(0.0, 1.0 - mass) * n
    big = exp >= 5
    chi2 = float((((obs - exp) ** 2)[big] / exp[big]).sum())
    po_, pe = obs[~big].sum(), exp[~big].sum()
    cells = int(big.sum())
    if pe >= 5:
        chi2 += (po_ - pe) ** 2 / pe
        cells += 1
    elif po_ > 0 and pe < 1e-3:
        chi2 += 1e9
    return float(stats.chi2.sf(chi2, cells - 1)), mass


MATS = {
    "diffuse": (po.material((0.8, 0.8, 0.8), specular_weight=0.0), 1.5),
    "conductor_r0.3": (po.material((0.95, 0.64, 0.54), metallic=1.0, roughness=0.3), 1.5),
    "conductor_r0.7": (po.material((0.95, 0.64, 0.54), metallic=1.0, roughness=0.7), 1.5),
    "glossy_diffuse_r0.3": (po.material((0.2, 0.4, 0.8), roughness=0.3), 1.5),
    "glossy_diffuse_r0.7": (po.material((0.2, 0.4, 0.8), roughness=0.7), 1.5),
    "dielectric_out_r0.3": (po.material((1, 1, 1), roughness=0.3, transmission=1.0), 1.5),
    "dielectric_in_r0.5": (po.material((1, 1, 1), roughness=0.5, transmission=1.0), 1 / 1.5),
    "mixture": (po.material((0.7, 0.5, 0.3), metallic=0.3, roughness=0.5, transmission=0.4), 1.5),
}


@pytest.mark.parametrize("name", sorted(MATS))
@pytest.mark.parametrize("theta", [30, 75])
def test_bsdf_sampling_matches_pdf(name, theta):
    mat, eta = MATS[name]
    p, mass = chi2_pvalue(mat, theta, eta)
    assert p > 1e-3, (name, theta, p)
    assert mass <= 1.0 + 2e-3


def test_chi2_detects_a_wrong_pdf():
    """the chi-square harness fails a sampler whose pdf is off: a conductor sampled at roughness 0.3 but evaluated
    at 0.35"""
    mat = po.material((0.95, 0.64, 0.54), metallic=1.0, roughness=0.3)
    wrong = po.material((0.95, 0.64, 0.54), metallic=1.0, roughness=0.35)
    wo = np.array([0.5, 0.0, math.sqrt(0.75)], np.float32)
    wi, w, pdf, ok = po.bsdf_sample_batch(mat, wo, 1.5, 200_000, 3)
    # re-run the test with the wrong pdf: patch the evaluation through a wrapper material
    orig = po.bsdf_pdf_batch
    try:
        po.bsdf_pdf_batch = lambda m, a, e, d, gd_energy=1: orig(wrong, a, e, d, gd_energy)
        p, _ = chi2_pvalue(mat, 30, 1.5)
    finally:
        po.bsdf_pdf_batch = orig
    assert p < 1e-6


@pytest.mark.parametrize("rough", [0.0, 0.1, 0.3, 0.5, 0.7, 1.0])
@pytest.mark.parametrize("theta", [1.0, 45.0, 80.0])
def test_white_furnace(rough, theta):
    """Monte Carlo albedo E[weight] of a non-absorbing conductor (F0 = 1), a lossless dielectric from both sides,
    and white glossy-diffuse with energy compensation (exactly 1), against the quadrature; within 1 %."""
    wo = np.array([math.sin(math.radians(theta)), 0.0, math.cos(math.radians(theta))], np.float32)
    delta = rough * rough < 1e-3
    cases = [("conductor", po.material((1, 1, 1), metallic=1.0, roughness=rough), 1.5, 0),
             ("dielectric_out", po.material((1, 1, 1), roughness=rough, transmission=1.0), 1.5, 1),
             ("dielectric_in", po.material((1, 1, 1), roughness=rough, transmission=1.0), 1 / 1.5, 1),
             ("glossy_diffuse", po.material((1, 1, 1), roughness=rough), 1.5, 2)]
    for name, mat, eta, kind in cases:
        wi, w, pdf, ok = po.bsdf_sample_batch(mat, wo, eta, 200_000, 11)
        lum = 0.2126 * w[:, 0] + 0.7152 * w[:, 1] + 0.0722 * w[:, 2]
        x = np.where(ok != 0, lum, 0.0).astype(np.float64)
        mc, se = x.mean(), x.std() / math.sqrt(len(x))
        ref = 1.0 if (delta or kind == 2) else po.lib().orc_quad_albedo(kind, rough, eta, wo.ctypes.data_as(po._FP))
        assert abs(mc - ref) <= 0.01 * ref + 4 * se, (name, rough, theta, mc, ref)
        if kind == 2:
            assert mc <= 1.0 + 4 * se


def test_fresnel_and_ggx_closed_forms():
    L = po.lib()
    for eta in (1.33, 1.5, 2.4, 1 / 1.5):
        assert L.orc_fresnel_dielectric(1.0, eta) == pytest.approx(((eta - 1) / (eta + 1)) ** 2, rel=1e-5)
    assert L.orc_fresnel_dielectric(0.1, 1 / 1.5) == 1.0            # total internal reflection
    f0, tint = np.array([0.9, 0.6, 0.3], np.float32), np.array([0.8, 0.9, 1.0], np.float32)
    out = np.zeros(3, np.float32)
    L.orc_fresnel_f82(po._f(f0), po._f(tint), 1.0, po._f(out))
    np.testing.assert_allclose(out, f0, rtol=1e-6)                # F(1) = F0
    L.orc_fresnel_f82(po._f(f0), po._f(tint), 1 / 7, po._f(out))
    schlick = f0 + (1 - f0) * (6 / 7) ** 5
    np.testing.assert_allclose(out, tint * schlick, rtol=1e-5)    # F(1/7) = tint * Schlick(1/7)
    one = np.ones(3, np.float32)
    for mu in (0.05, 0.3, 0.8):
        L.orc_fresnel_f82(po._f(f0), po._f(one), mu, po._f(out))
        np.testing.assert_allclose(out, f0 + (1 - f0) * (1 - mu) ** 5, rtol=1e-5)   # tint 1: plain Schlick
    for a in (0.05, 0.3, 0.8):   # projected-area normalisation: 2 pi * integral_0^1 D(c) c dc = 1
        n = 20_000
        u = (np.arange(n) + 0.5) / n
        c = 1.0 - u ** 4                                           # dense near c = 1, where D peaks
        D = np.array([L.orc_ggx_d(float(x), a * a) for x in c])
        integral = 2 * np.pi * np.sum(D * c * 4 * u ** 3) / n
        assert integral == pytest.approx(1.0, rel=2e-3)
    assert L.orc_ggx_g2(1.0, 1.0, 0.25) == pytest.approx(1.0)      # no masking at normal incidence


def _env_scene(tex, w, h, uniform=0):
    sc = po.Scene(2, 0, np.zeros((1, 16), np.float32), np.zeros((1, 12), np.float32), np.zeros((1, 12), np.float32),
                  np.ones((1, 4), np.float32))
    sc.set_full(po.material(), np.zeros(1, np.int32), env_tex=tex, env_w=w, env_h=h, env_uniform=uniform)
    return sc


def test_environment_distribution():
    """the luminance sampling's pdf integrates to 1 over the sphere and matches the histogram of its samples"""
    w, h = 32, 16
    rng = np.random.default_rng(5)
    rgb = np.exp(rng.normal(0, 1.5, (h, w, 1))).astype(np.float32) * np.array([1.0, 0.9, 0.8], np.float32)
    tex, marg, cond = po.env_build(rgb, 2.0)
    sc = _env_scene(tex, w, h)
    sub = 8
    d = _dirs_grid(64, 128, 1)
    pdf = np.array([po.lib().orc_env_pdf(po.ctypes.byref(sc.c), po._f(np.ascontiguousarray(x))) for x in d])
    assert (4 * np.pi * pdf.mean()) == pytest.approx(1.0, rel=2e-3)
    n = 400_000
    dirs, p = np.zeros((n, 3), np.float32), np.zeros(n, np.float32)
    po.lib().orc_env_sample_batch(po.ctypes.byref(sc.c), n, 9, po._f(dirs), po._f(p))
    # histogram over texels vs expected probability = luminance * solid angle share
    u = np.mod(0.5 - np.arctan2(dirs[:, 1], dirs[:, 0]) / (2 * np.pi), 1.0)
    v = np.arccos(np.clip(dirs[:, 2], -1, 1)) / np.pi
    x = np.minimum((u * w).astype(int), w - 1); y = np.minimum((v * h).astype(int), h - 1)
    obs = np.bincount(y * w + x, minlength=w * h)
    lum = (0.2126 * tex[:, 0] + 0.7152 * tex[:, 1] + 0.0722 * tex[:, 2]).reshape(h, w).astype(np.float64)
    st = np.sin(np.pi * (np.arange(h) + 0.5) / h)[:, None]
    prob = (lum * st / (lum * st).sum()).reshape(-1)
    exp = prob * n
    m = exp >= 5
    chi2 = (((obs - exp) ** 2)[m] / exp[m]).sum()
    assert stats.chi2.sf(chi2, m.sum() - 1) > 1e-3
    assert np.all(p > 0)


def test_product_light_tables_equal_the_oracle():
    """emitter CDF (mesh order) and environment CDFs of the product's host setup equal the oracle's bit for bit"""
    s = cwp.Setup(["--gen", "terrain", "--gen_grid", "30", "--gen_boxes", "40", "--gen_materials", "2", "--gen_env", "64,32,2,4",
                   "--width", "16", "--height", "16"])
    info = s.info
    assert info["full"] == 1 and info["num_emitters"] > 0 and info["env_w"] == 64
    sc, b = po.scene_from_setup(s)
    np.testing.assert_array_equal(s.array(cwp.ARR_EM_CDF), sc.em_cdf)
    # the same emitters (leaf positions differ between the two BVHs; compare through the mesh triangle id)
    np.testing.assert_array_equal(s.array(cwp.ARR_ORDER)[s.array(cwp.ARR_EM_TRI)], np.asarray(b["order"])[sc.em_tri])
    np.testing.assert_array_equal(s.array(cwp.ARR_ENV_MARG), sc.env_marg)
    np.testing.assert_array_equal(s.array(cwp.ARR_ENV_COND), sc.env_cond.reshape(-1))
    # probabilities: sum over emitters of pdf * area = 1
    c = np.asarray(b["corners"]).reshape(-1, 3, 4)[sc.em_tri]
    area = 0.5 * np.linalg.norm(np.cross(c[:, 1, :3] - c[:, 0, :3], c[:, 2, :3] - c[:, 0, :3]), axis=1)
    assert float((sc.tri_light_pdf[sc.em_tri] * area).sum()) == pytest.approx(1.0, rel=1e-5)


def _plane_setup(tmp_path, extra):
    obj = tmp_path / "plane.obj"
    obj.write_text("v -1000 -1000 0\nv 1000 -1000 0\nv 1000 1000 0\nv -1000 1000 0\nf 1 2 3\nf 1 3 4\n")
    return cwp.Setup(["--mesh", str(obj), "--c2w", "1,0,0,0,0,0.7071068,-0.7071068,-2,0,0.7071068,0.7071068,1,0,0,0,1",
                      "--width", "32", "--height", "32", "--albedo", "0.6", "0.6", "0.6", "--sun_rgb", "0", "0", "0",
                      "--filter", "box", "--integrator", "full", "--bounces", "3"] + extra)


def _hemisphere_mean(tex, w, h):
    """(1 / pi) * integral over the upper hemisphere of L(w) cos(theta) dw, by 8 x 8 sub-samples per texel"""
    L = (0.2126 * tex[:, 0] + 0.7152 * tex[:, 1] + 0.0722 * tex[:, 2]).reshape(h, w).astype(np.float64)
    s = 8
    v = (np.arange(h * s) + 0.5) / (h * s)
    th = v * np.pi
    cos_t = np.cos(th)
    dA = (np.pi / (h * s)) * (2 * np.pi / (w * s)) * np.sin(th)
    rowL = np.repeat(L, s, axis=0).mean(axis=1)   # each texel row's mean over columns (columns cover phi uniformly)
    up = cos_t > 0
    return float((rowL[up] * cos_t[up] * dA[up] * (w * s)).sum() / np.pi)


@pytest.mark.parametrize("mode", ["mis", "nee", "bsdf"])
def test_plane_under_environment_closed_form(tmp_path, mode):
    a, spp = 0.6, 64
    w, h = 32, 16
    s = _plane_setup(tmp_path, ["--gen_env", "%d,%d,1.0,3" % (w, h), "--light_sampling", mode])
    tex = s.array(cwp.ARR_ENV_TEX)
    sc, _ = po.scene_from_setup(s)
    acc, _ = sc.render(s.params(), spp, batch=spp, bounces=3)
    img = acc[..., :3] / spp
    lum = 0.2126 * img[..., 0] + 0.7152 * img[..., 1] + 0.0722 * img[..., 2]
    expect = a * _hemisphere_mean(tex, w, h)
    plane = lum[16:]                               # the lower half of the image sees the plane
    se = plane.std() / math.sqrt(plane.size)
    assert plane.mean() == pytest.approx(expect, abs=4 * se + 0.01 * expect), (mode, plane.mean(), expect)


def test_closed_emissive_box(tmp_path):
    """all six walls emit E and reflect with Lambert albedo a; camera inside: E (1 - a^(B+2)) / (1 - a) per pixel"""
    obj = tmp_path / "box.obj"
    v = [(x, y, z) for x in (-1, 1) for y in (-1, 1) for z in (-1, 1)]
    faces = [(1, 3, 4, 2), (5, 6, 8, 7), (1, 2, 6, 5), (3, 7, 8, 4), (1, 5, 7, 3), (2, 4, 8, 6)]
    lines = ["v %d %d %d" % p for p in v]
    for q in faces:
        lines.append("f %d %d %d" % (q[0], q[1], q[2])); lines.append("f %d %d %d" % (q[0], q[2], q[3]))
    obj.write_text("\n".join(lines) + "\n")
    a, E, B, spp = 0.5, 1.0, 3, 64
    mats = tmp_path / "mats.json"
    mats.write_text('{"materials": [{"base_color": [%g, %g, %g], "specular_weight": 0, "roughness": 1,'
                    ' "emission_color": [1, 1, 1], "emission_strength": %g}]}' % (a, a, a, E))
    for mode in ("mis", "nee", "bsdf"):
        s = cwp.Setup(["--mesh", str(obj), "--materials", str(mats), "--c2w", "1,0,0,0.1,0,1,0,-0.2,0,0,1,0.05,0,0,0,1",
                       "--width", "24", "--height", "24", "--sun_rgb", "0", "0", "0", "--filter", "box", "--bounces", str(B),
                       "--light_sampling", mode, "--hfov", "90"])
        assert s.info["full"] == 1
        sc, _ = po.scene_from_setup(s)
        acc, _ = sc.render(s.params(), spp, batch=spp, bounces=B)
        img = acc[..., :3] / spp
        expect = E * (1 - a ** (B + 2)) / (1 - a)
        se = img[..., 0].std() / math.sqrt(img[..., 0].size)
        assert img[..., 0].mean() == pytest.approx(expect, abs=4 * se + 2e-3), (mode, img[..., 0].mean(), expect)


def _inst_oracle(s, frame):
    i = s.info
    sc = po.Scene(i["variant"], i["tri_test"], s.array(cwp.ARR_NODES), s.array(cwp.ARR_WOOP), s.array(cwp.ARR_CORNERS),
                  s.array(cwp.ARR_ALBEDO))
    sc.set_instances(s.instances(frame), s.array(cwp.ARR_MESH_BOXES))
    sc.set_full(s.array(cwp.ARR_MATERIALS), s.array(cwp.ARR_ORDER), i["sun_on"], i["light_mode"], i["rr"], i["rr_depth"],
                i["gd_energy"], i["env_uniform"], s.array(cwp.ARR_ENV_TEX) if i["env_w"] else None, i["env_w"], i["env_h"],
                i["env_rot_deg"])
    sc.set_inst_emitters(s.array(cwp.ARR_MESH_TRIS), s.array(cwp.ARR_ORDER))
    return sc


def test_instanced_emitters_equal_the_oracle():
    """per-frame world emitters, their CDF and the per-material pdf of the product's host equal the oracle's; the
    light pdf integrates to one over the emitters"""
    s = cwp.Setup(["--gen", "instances", "--gen_small", "1", "--gen_instances", "30", "--gen_materials", "2",
                   "--width", "16", "--height", "16"])
    assert s.info["full"] == 1 and s.info["num_emitters"] > 0
    for f in (0, 5):
        em, cdf, mlp = s.emitters(f)
        sc = _inst_oracle(s, f)
        np.testing.assert_array_equal(em, sc.em_tris[:len(em)])
        np.testing.assert_array_equal(cdf, sc.em_cdf)
        np.testing.assert_array_equal(mlp, sc.mlp)
        area = 0.5 * np.linalg.norm(np.cross(em[:, 1, :3], em[:, 2, :3]), axis=1)
        mats = em[:, 1, 3].view(np.int32)
        assert float((mlp[mats] * area).sum()) == pytest.approx(1.0, rel=1e-5)


def test_instanced_full_render_equals_flattened_oracle():
    """the oracle's two-level full-integrator render equals its render of the flattened scene (same world
    triangles, same emitter CDF in instance order) to >= 60 dB: only the rounding of object-space intersection and
    transformed normals differs"""
    base = ["--gen", "instances", "--gen_small", "1", "--gen_instances", "30", "--gen_materials", "2", "--width", "32",
            "--height", "24", "--bounces", "3", "--gen_env", "32,16,1,1"]
    s = cwp.Setup(base)
    sc = _inst_oracle(s, 0)
    a, _ = sc.render(s.params(), 16, batch=16, bounces=3)
    f = cwp.Setup(base + ["--flatten", "1"])
    fc, _ = po.scene_from_setup(f)
    b, _ = fc.render(f.params(), 16, batch=16, bounces=3)
    mse = float(np.mean((np.clip(a[..., :3] / 16, 0, 1) - np.clip(b[..., :3] / 16, 0, 1)) ** 2))
    assert 10 * math.log10(1.0 / max(mse, 1e-12)) >= 60.0
