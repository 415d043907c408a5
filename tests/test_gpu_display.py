"""GPU parity of the display-tail stages against the oracle, through the C ABI.

HighlightReconstruction, Vignette and BaselineExposureCompensation use only
+-*/ and sqrt (correctly rounded on both sides): bit-exact.  Contrast (powf),
LCH and SigmoidToneMap (cbrtf, atan2f, sinf, cosf, powf) evaluate transcendentals
whose device and glibc results may differ by an ulp: max abs error <= 2e-5 on
values of order 1.
"""
import struct

import numpy as np
import pytest

import imops_amd as im
import oracle

pytestmark = pytest.mark.gpu
TOL = 2e-5


def _scene(w, h, seed, lo=0.0, hi=1.2):
    rng = np.random.default_rng(seed)
    return (im.synth_scene_rgb(w, h, seed) * np.float32(hi - lo) + np.float32(lo)
            + rng.normal(0, 0.05, (h, w, 3)).astype(np.float32)).astype(np.float32)


def _img(ctx, rgb, cs=im.CS_ACESCG):
    return im.Image(ctx).set_rgb(rgb, cs)


def _exact(got, want):
    bad = np.count_nonzero(got.view(np.uint32) != want.view(np.uint32))
    assert bad == 0, f"{bad} mismatching values, max abs {np.abs(got - want).max()}"


def test_contrast(gpu_ctx):
    x = np.abs(_scene(97, 61, 1))
    got = _img(gpu_ctx, x).contrast(1.35).rgb()
    assert np.abs(got - oracle.contrast(x, 1.35)).max() <= TOL


def test_highlight_reconstruction(gpu_ctx):
    x = _scene(97, 61, 2)
    wb = np.array([2.1, 1.3, 1.7, 1.3], np.float32)
    got = _img(gpu_ctx, x).set_wb_coeffs(wb).highlight_reconstruction().rgb()
    _exact(got, oracle.highlight(x, np.float32(1.3)))


@pytest.mark.parametrize("cs", [im.CS_SRGB, im.CS_ACESCG, im.CS_LINEAR_SRGB])
def test_lch(gpu_ctx, cs):
    x = np.clip(_scene(80, 50, 3, 0.02, 0.9), 0.001, 1.0)
    got = _img(gpu_ctx, x, cs).lch(1.1, 0.7, 1.05).rgb()
    assert np.abs(got - oracle.lch(x, cs, 1.1, 0.7, 1.05)).max() <= TOL


def test_sigmoid_tone_map(gpu_ctx):
    x = _scene(120, 70, 4, -0.05, 3.0)  # includes negative (out-of-gamut) and bright values
    got = _img(gpu_ctx, x).sigmoid_tone_map().rgb()
    assert np.abs(got - oracle.sigmoid(x)).max() <= TOL


def test_baseline_exposure(gpu_ctx):
    x = _scene(64, 48, 5)
    img = _img(gpu_ctx, x).set_shutter_seconds(1 / 125).set_exposure_metadata(0.35, 5.6, 200.0)
    _exact(img.baseline_exposure().rgb(), oracle.baseline(x, 0.35, np.float32(1 / 125), 5.6, 200.0))


def _opcodes(k, cx, cy):
    body = struct.pack(">7d", *k, cx, cy)
    skip = struct.pack(">IIII", 1, 1, 0, 4) + b"\0" * 4
    return struct.pack(">I", 2) + skip + struct.pack(">IIII", 3, 1, 0, len(body)) + body


@pytest.mark.parametrize("w,h", [(97, 61), (300, 2), (1, 9)])
def test_vignette(gpu_ctx, w, h):
    x = _scene(w, h, 6)
    ops = _opcodes((-0.4, 0.15, 0.02, 0.0, -0.01), 0.47, 0.55)
    got = _img(gpu_ctx, x).set_opcode_list3(ops).vignette(0.9).rgb()
    _exact(got, oracle.vignette(x, ops, 0.9))
    # no opcode list: unchanged
    _exact(_img(gpu_ctx, x).vignette(0.9).rgb(), x)


def test_display_tail_pipeline(gpu_ctx):
    x = np.abs(_scene(90, 60, 7))
    ops = _opcodes((-0.3, 0.1, 0.0, 0.0, 0.0), 0.5, 0.5)
    cfg = im.parse_config("""
[[pipeline_modules]]
name = "BaselineExposureCompensation"
[[pipeline_modules]]
name = "Vignette"
strength = 0.7
[[pipeline_modules]]
name = "HighlightReconstruction"
[[pipeline_modules]]
name = "SigmoidToneMap"
[[pipeline_modules]]
name = "Contrast"
c = 1.1
[[pipeline_modules]]
name = "LCH"
lc = 1.0
cc = 1.2
hc = 1.0
""")
    wb = np.array([1.9, 250.0, 1.5, 250.0], np.float32)
    img = _img(gpu_ctx, x).set_wb_coeffs(wb).set_shutter_seconds(1 / 60).set_exposure_metadata(0.0, 4.0, 100.0)
    img.set_opcode_list3(ops)
    im.run_pixel_pipeline(img, cfg)
    want = oracle.baseline(x, 0.0, np.float32(1 / 60), 4.0, 100.0)
    want = oracle.vignette(want, ops, 0.7)
    want = oracle.highlight(want, 250.0)
    want = oracle.contrast(oracle.sigmoid(want), 1.1)
    want = oracle.lch(want, oracle.CS_ACESCG, 1.0, 1.2, 1.0)
    got = img.rgb()
    assert np.array_equal(np.isnan(got), np.isnan(want))
    assert np.nanmax(np.abs(got - want)) <= 5e-5
