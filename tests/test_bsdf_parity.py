"""The library's material response (core/bsdf.h, the code the full-integrator kernels run, called here on the host
through cwp_bsdf_eval / cwp_bsdf_sample) against the oracle's restatement of the reference's bsdfEval / bsdfSample
(src/render/bsdf.h:196-220). The library evaluates the three bases with shared microfacet terms and the oracle lobe by
lobe, so values agree to float rounding, not bit for bit: relative 2e-5 on f and pdf, the same accept / reject
decision and the same direction on all but borderline samples."""
import ctypes

import numpy as np
import pytest

import cwp
import pyoracle as po


def _mat(base, metallic, spec, rough, ior, ws, trans):
    return np.array(list(base) + [metallic] + list(spec) + [rough] + [0, 0, 0] + [ior, ws, trans, 0, 0], np.float32)


MATERIALS = {
    "lambert": _mat((0.7, 0.5, 0.3), 0, (1, 1, 1), 1.0, 1.5, 0.0, 0),
    "glossy-diffuse": _mat((0.8, 0.2, 0.1), 0, (1, 1, 1), 0.35, 1.5, 1.0, 0),
    "glossy-smooth": _mat((0.2, 0.6, 0.3), 0, (1, 1, 1), 0.02, 1.45, 0.8, 0),
    "conductor": _mat((0.95, 0.64, 0.54), 1, (0.9, 0.8, 0.7), 0.3, 1.5, 1.0, 0),
    "conductor-smooth": _mat((0.9, 0.9, 0.9), 1, (1, 1, 1), 0.01, 1.5, 1.0, 0),
    "glass-rough": _mat((1, 0.9, 0.8), 0, (1, 1, 1), 0.25, 1.5, 1.0, 1),
    "glass-smooth": _mat((1, 1, 1), 0, (1, 1, 1), 0.0, 1.5, 1.0, 1),
    "blend": _mat((0.6, 0.5, 0.4), 0.4, (0.8, 0.9, 1.0), 0.45, 1.6, 0.7, 0.5),
}


def _dirs(rng, n, upper):
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    if upper:
        d[:, 2] = np.abs(d[:, 2]) + 1e-3
        d /= np.linalg.norm(d, axis=1, keepdims=True)
    return d.astype(np.float32)


def _close(a, b, rtol=2e-5):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    scale = max(np.max(np.abs(b)), 1e-30)
    return bool(np.all(np.abs(a - b) <= rtol * np.maximum(np.abs(b), 1e-3 * scale) + 1e-30))


@pytest.mark.parametrize("name", sorted(MATERIALS))
@pytest.mark.parametrize("eta", [1.5, 1.0 / 1.5])
def test_eval_matches_oracle(name, eta):
    m = MATERIALS[name]
    rng = np.random.default_rng(7)
    wos, wis = _dirs(rng, 300, True), _dirs(rng, 300, False)
    ux = rng.uniform(size=(300, 2)).astype(np.float32)
    for k in range(300):
        for gde in (0, 1):
            f, p = cwp.bsdf_eval(m, wos[k], wis[k], eta, gde, float(ux[k, 0]), float(ux[k, 1]))
            fo, pref = po.bsdf_eval(m, wos[k], wis[k], eta, gde, float(ux[k, 0]), float(ux[k, 1]))
            assert _close(f, fo) and _close([p], [pref]), (name, eta, gde, k, f, fo, p, pref)


def _oracle_sample(m, wo, eta, gde, ux1, ux2, u, u1, u2):
    f = lambda a: np.ascontiguousarray(a, np.float32).ctypes.data_as(ctypes.POINTER(ctypes.c_float))
    wi, w, p = np.zeros(3, np.float32), np.zeros(3, np.float32), np.zeros(1, np.float32)
    ok = po.lib().orc_bsdf_sample(f(m), f(wo), eta, gde, ux1, ux2, u, u1, u2,
                                  wi.ctypes.data_as(ctypes.POINTER(ctypes.c_float)),
                                  w.ctypes.data_as(ctypes.POINTER(ctypes.c_float)),
                                  p.ctypes.data_as(ctypes.POINTER(ctypes.c_float)))
    return bool(ok), wi, w, float(p[0])


@pytest.mark.parametrize("name", sorted(MATERIALS))
def test_sample_matches_oracle(name):
    m = MATERIALS[name]
    rng = np.random.default_rng(11)
    n = 2000
    wos = _dirs(rng, n, True)
    us = rng.uniform(size=(n, 5)).astype(np.float32)
    differ = 0
    for k in range(n):
        eta = 1.5 if k % 2 else 1.0 / 1.5
        gde = k % 3 != 0
        args = (float(us[k, 3]), float(us[k, 4]), float(us[k, 0]), float(us[k, 1]), float(us[k, 2]))
        ok, wi, w, p = cwp.bsdf_sample(m, wos[k], eta, args[2], args[3], args[4], int(gde), args[0], args[1])
        ok2, wi2, w2, p2 = _oracle_sample(m, wos[k], eta, int(gde), *args)
        if ok != ok2 or (p == 0) != (p2 == 0):
            differ += 1
            continue
        if not ok:
            continue
        assert np.allclose(wi, wi2, atol=2e-5), (name, k, wi, wi2)
        assert _close(w, w2, 1e-4) and _close([p], [p2], 1e-4), (name, k, w, w2, p, p2)
    assert differ <= n // 500, (name, differ)


def test_a_wrong_index_ratio_is_caught():
    """The comparison is sharp: the oracle evaluated with the dielectric base's index ratio off by 1 % disagrees."""
    m = MATERIALS["glass-rough"]
    rng = np.random.default_rng(3)
    wo, wi = _dirs(rng, 1, True)[0], _dirs(rng, 1, True)[0]
    f, p = cwp.bsdf_eval(m, wo, wi, 1.5, 1)
    fo, pref = po.bsdf_eval(m, wo, wi, 1.5 * 1.01, 1)
    assert not (_close(f, fo) and _close([p], [pref]))
