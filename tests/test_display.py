"""The viewer's display transform on the host: the oracle's restatement (src/app/display.cu:9-33) against the sRGB
standard's 8-bit values (IEC 61966-2-1) and the tone curves' closed forms, the step table the kernel searches
against the oracle's encode for every float in [0, 1], the library's host encode, and the headless viewer's option
errors (which it reports before it touches a GPU)."""
import math
import os
import subprocess

import numpy as np
import pytest

import cwp
import pyoracle as po

VIEWER = os.path.join(os.path.dirname(cwp.__file__), "viewer")


# ---------------------------------------------------------------------------------------- the 8-bit sRGB encode
@pytest.mark.parametrize("lin,byte", [
    (0.0, 0), (1.0, 255),
    (0.5, 188),              # linear 0.5 -> sRGB 0.7354
    (0.18, 118),             # 18 % grey -> sRGB 0.4614
    (0.2158, 128),           # the linear value of sRGB 128 (0.5019)
    (0.25, 137),             # sRGB 0.5371
    (0.0031308, 10),         # the end of the linear toe: 12.92 x = 0.04045
    (0.001, 3),              # on the toe: 12.92 * 0.001 * 255 = 3.29
    (2.0, 255), (-0.5, 0), (float("inf"), 255), (float("-inf"), 0), (float("nan"), 0),
])
def test_oracle_encode_matches_the_srgb_standard(lin, byte):
    assert po.srgb_encode8(lin) == byte
    assert cwp.srgb_encode8(lin) == byte


def test_encode_is_monotone_and_hits_every_byte():
    x = np.linspace(0.0, 1.0, 200001, dtype=np.float32)
    b = np.array([po.srgb_encode8(float(v)) for v in x[::7]])
    assert np.all(np.diff(b) >= 0)
    assert set(b.tolist()) == set(range(256))


def test_step_table_equals_the_encode_for_every_float_in_the_unit_interval():
    """the kernel's search over 255 step positions gives the host encode's byte for all 1,065,353,217 floats in
    [0, 1] (the encode clips everything else to these)"""
    t = cwp.srgb_steps()
    assert np.isneginf(t[0]) and np.all(np.diff(t[1:]) > 0)
    assert t[1] > 0 and t[255] < 1
    for k in (1, 2, 10, 128, 188, 255):        # each step is the first float of its byte
        assert po.srgb_encode8(float(t[k])) == k
        assert po.srgb_encode8(float(np.nextafter(t[k], np.float32(0)))) == k - 1
    assert po.srgb_steps_check(t, 0.0, 1.0, 8) == 0


def test_a_wrong_step_table_is_caught():
    t = cwp.srgb_steps()
    t[100] = np.nextafter(t[100], np.float32(1))   # one step a float late: that float encodes one lower
    assert po.srgb_steps_check(t, 0.0, 1.0, 8) == 1


# ------------------------------------------------------------------------------------------ the display transform
def acc_of(rgb, spp):
    return np.array([[c * spp for c in rgb] + [0.0]], np.float32)


def test_clip_mode_at_exposure_zero_is_the_png_encode_of_the_mean():
    rng = np.random.default_rng(5)
    a = rng.uniform(-0.5, 3.0, (2000, 4)).astype(np.float32)
    for spp in (1, 3, 64):
        d = po.display(a, spp, 0.0, 0)
        mean = a / np.float32(spp)
        want = np.array([[po.srgb_encode8(float(v)) for v in row[:3]] for row in mean], np.uint8)
        assert np.array_equal(d[:, :3], want)
        assert np.all(d[:, 3] == 255)


def test_spp_zero_is_read_as_one():
    a = np.array([[0.25, 0.5, 2.0, 0.0]], np.float32)
    assert np.array_equal(po.display(a, 0), po.display(a, 1))
    assert po.display(a, 1)[0].tolist() == [137, 188, 255, 255]


def test_exposure_scales_by_powers_of_two():
    rng = np.random.default_rng(6)
    a = rng.uniform(0, 1.5, (3000, 4)).astype(np.float32)
    for mode in (0, 1, 2):
        assert np.array_equal(po.display(a, 4, 1.0, mode), po.display(a * 2, 4, 0.0, mode))
        assert np.array_equal(po.display(a, 4, -2.0, mode), po.display(a / 4, 4, 0.0, mode))


def test_reinhard_closed_forms():
    # x / (1 + x): 1 -> 1/2, 3 -> 3/4, 1/3 -> 1/4; negative means clamp to 0 before the curve
    assert po.display(acc_of([1.0, 3.0, 1.0 / 3.0], 2), 2, 0.0, 1)[0, :3].tolist() == [
        po.srgb_encode8(0.5), po.srgb_encode8(0.75), po.srgb_encode8(0.25)]
    assert po.display(acc_of([-1.0, 0.0, 1e30], 1), 1, 0.0, 1)[0, :3].tolist() == [0, 0, 255]


def test_aces_limits_and_monotonicity():
    # Narkowicz's fit after the 0.6 scale: f(0) = 0, f -> 2.51 / 2.43 > 1 (clipped), increasing
    x = np.linspace(0, 20, 4001, dtype=np.float32)
    a = np.stack([x, x, x, np.zeros_like(x)], 1)
    d = po.display(a, 1, 0.0, 2)[:, 0].astype(int)
    assert d[0] == 0 and d[-1] == 255 and np.all(np.diff(d) >= 0)
    # at x = 1 / 0.6 the scaled value is 1: (2.51 + 0.03) / (2.43 + 0.59 + 0.14) = 0.80380
    v = np.float32(1.0) / np.float32(0.6)
    assert abs(po.display(acc_of([v, v, v], 1), 1, 0.0, 2)[0, 0] - po.srgb_encode8(2.54 / 3.16)) <= 1


def test_tone_curves_differ_from_clip_only_above_zero():
    a = np.array([[0.0, 0.0, 0.0, 0.0], [-2.0, -1.0, -0.1, 0.0]], np.float32)
    for mode in (1, 2):
        assert np.array_equal(po.display(a, 1, 0.0, mode), po.display(a, 1, 0.0, 0))


# ------------------------------------------------------------------------------------------- headless viewer CLI
@pytest.mark.skipif(not os.path.exists(VIEWER), reason="viewer not built")
@pytest.mark.parametrize("args,msg", [
    (["--interop", "1"], "interactive window"),
    (["--win_width", "800"], "interactive window"),
    (["--screenshot_window", "w.png"], "interactive window"),
    (["--tonemap", "filmic"], "--tonemap must be"),
    (["--frames", "0"], "--frames must be"),
    (["--passes", "x"], "--passes must be"),
    (["--exposure", "nan"], "--exposure must be"),
    (["--bogus"], "bogus"),
])
def test_viewer_option_errors(args, msg):
    r = subprocess.run([VIEWER, "--gen", "boxes", "--width", "8", "--height", "8"] + args, capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 1
    assert msg in r.stderr
