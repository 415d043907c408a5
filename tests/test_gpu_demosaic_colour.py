"""GPU parity: raw ingest, AMaZE / Markesteijn demosaic and the colour stages
against the oracle, through the C ABI.

Bar: bit-exact for ingest and demosaic (pure +-*/ f32 arithmetic in the
reference's order); colour stages that call powf (sRGB OETF/EOTF, Gamma)
within 2e-6 absolute (device vs glibc powf differ by at most an ulp or two),
pure matrix stages bit-exact.
"""
import numpy as np
import pytest

import imops_amd as im
import oracle

pytestmark = pytest.mark.gpu

BLACK, WHITE = 600.0, 16383.0

# random diagonally dominant XYZ->camera matrix with cond < 10 (seeded)
def _calibration(seed=7):
    rng = np.random.default_rng(seed)
    while True:
        m = rng.uniform(-0.25, 0.25, (3, 3)) + np.diag(rng.uniform(0.8, 1.3, 3))
        if np.linalg.cond(m) < 10:
            return m.astype(np.float32)


def _ingest(ctx, w, h, cfa="RGGB", frame=0, crop=None):
    raw = im.synth_raw_u16(w, h, cfa, frame)
    img = im.Image(ctx).set_raw_u16(raw, cfa, BLACK, WHITE, crop=crop)
    ref = oracle.normalize_raw(raw, BLACK, WHITE, crop=crop)
    return raw, img, ref


def _mismatch(a, b):
    return int(np.count_nonzero(a.view(np.uint32) != b.view(np.uint32)))


@pytest.mark.parametrize("w,h", [(1024, 1024), (256, 192), (130, 113), (400, 288), (112, 208), (40, 34)])
def test_amaze_bit_exact(gpu_ctx, w, h):
    print(f"amaze {w}x{h}", flush=True)
    _, img, plane = _ingest(gpu_ctx, w, h)
    img.demosaic("amaze")
    got = img.rgb()
    want = oracle.amaze(plane, "RGGB")
    assert _mismatch(got, want) == 0, f"max abs {np.abs(got - want).max()}"
    assert img.color_space == im.CS_NONE


@pytest.mark.parametrize("cfa,crop", [("BGGR", None), ("GRBG", (1, 0, 320, 288)), ("RGGB", (3, 5, 288, 320))])
def test_amaze_patterns_and_crop(gpu_ctx, cfa, crop):
    print(f"amaze {cfa} crop={crop}", flush=True)
    raw, img, plane = _ingest(gpu_ctx, 323, 326, cfa, frame=3, crop=crop or (0, 0, 320, 320))
    img.demosaic("amaze")
    c = crop or (0, 0, 320, 320)
    want = oracle.amaze(plane, cfa, crop_origin=(c[0], c[1]))
    assert _mismatch(img.rgb(), want) == 0


def test_amaze_uniform_closed_form(gpu_ctx):
    # demosaic.rs:1418-1443 checks |out-0.5| < 0.2; the algorithm is exact on a flat field
    img = im.Image(gpu_ctx).set_raw_f32(np.full((160, 160), 0.5, np.float32), "RGGB")
    img.demosaic("amaze")
    assert np.all(img.rgb() == np.float32(0.5))


def test_amaze_rejects_unsupported(gpu_ctx):
    img = im.Image(gpu_ctx).set_raw_f32(np.zeros((200, 200), np.float32), "RGGB")  # 200 % 96 = 8
    with pytest.raises(im.ImopsError):
        img.demosaic("amaze")
    xt = "GGRGGB" "GGBGGR" "BRGRBG" "GGBGGR" "GGRGGB" "RBGBRG"
    img2 = im.Image(gpu_ctx).set_raw_f32(np.zeros((192, 192), np.float32), xt)
    with pytest.raises(im.ImopsError):
        img2.demosaic("amaze")


XTRANS = "GGRGGB" "GGBGGR" "BRGRBG" "GGBGGR" "GGRGGB" "RBGBRG"


@pytest.mark.parametrize("cfa,w,h", [("RGGB", 257, 131), (XTRANS, 624, 416), (XTRANS, 17, 9)])
def test_markesteijn_bit_exact(gpu_ctx, cfa, w, h):
    print(f"markesteijn {w}x{h}", flush=True)
    _, img, plane = _ingest(gpu_ctx, w, h, cfa, frame=5)
    img.demosaic("markesteijn")
    want = oracle.markesteijn(plane, cfa)
    assert _mismatch(img.rgb(), want) == 0


@pytest.mark.parametrize("algorithm,factor", [("fast", 2), ("superfast", 4), ("supersuperfast", 8)])
@pytest.mark.parametrize("cfa,w,h", [("RGGB", 257, 131), ("BGGR", 64, 48), (XTRANS, 99, 50)])
def test_preview_demosaic_bit_exact(gpu_ctx, algorithm, factor, cfa, w, h):
    _, img, plane = _ingest(gpu_ctx, w, h, cfa, frame=7)
    img.demosaic(algorithm)
    want = oracle.preview_demosaic(plane, cfa, factor)
    assert img.info()[:2] == (w // factor, h // factor)
    assert _mismatch(img.rgb(), want) == 0


def test_preview_demosaic_crop_and_pipeline(gpu_ctx):
    raw = im.synth_raw_u16(120, 90, "RGGB", 11)
    img = im.Image(gpu_ctx).set_raw_u16(raw, "RGGB", BLACK, WHITE, crop=(3, 2, 100, 80))
    im.run_pixel_pipeline(img, im.parse_config('[[pipeline_modules]]\nname = "Demosaic"\nalgorithm = "fast"\n'))
    plane = oracle.normalize_raw(raw, BLACK, WHITE, crop=(3, 2, 100, 80))
    assert _mismatch(img.rgb(), oracle.preview_demosaic(plane, "RGGB", 2, crop=(3, 2))) == 0


def test_ingest_device_buffer_and_crop(gpu_ctx):
    import ctypes
    raw = im.synth_raw_u16(200, 100, "RGGB", 9)
    # normalisation of a crop window must match the oracle bit for bit
    img = im.Image(gpu_ctx).set_raw_u16(raw, "RGGB", BLACK, WHITE, crop=(5, 7, 150, 80))
    img.demosaic("markesteijn")
    plane = oracle.normalize_raw(raw, BLACK, WHITE, crop=(5, 7, 150, 80))
    assert _mismatch(img.rgb(), oracle.markesteijn(plane, "RGGB", crop_origin=(5, 7))) == 0


@pytest.mark.parametrize("target", [im.CS_SRGB, im.CS_ACESCG, im.CS_XYZD65, im.CS_LINEAR_SRGB])
def test_camera_cst(gpu_ctx, target):
    rgb = im.synth_scene_rgb(300, 200, 4)
    cal = _calibration()
    img = im.Image(gpu_ctx).set_rgb(rgb, im.CS_NONE).set_calibration_matrix(cal.ravel())
    img.cst(target)
    got = img.rgb()
    want = oracle.camera_cst(rgb, cal.ravel(), target)
    if target == im.CS_SRGB:
        assert np.abs(got - want).max() <= 2e-6
    else:
        assert _mismatch(got, want) == 0
    assert img.color_space == target


def test_cst_chain_and_scalars(gpu_ctx):
    rgb = im.synth_scene_rgb(128, 96, 2) * 3.0
    img = im.Image(gpu_ctx).set_rgb(rgb.astype(np.float32), im.CS_ACESCG)
    img.exp(1.5)
    img.cst("Srgb")
    img.gamma(2.4)
    want = oracle.gamma(oracle.cst(oracle.exp(rgb.astype(np.float32), 1.5), oracle.CS_ACESCG, oracle.CS_SRGB), 2.4)
    assert np.abs(img.rgb() - want).max() <= 2e-6
    img2 = im.Image(gpu_ctx).set_rgb(rgb.astype(np.float32), im.CS_NONE)
    img2.cfa_coeffs([2.0, 1.0, 1.5, 1.0])
    assert _mismatch(img2.rgb(), oracle.cfa_coeffs(rgb.astype(np.float32), [2.0, 1.0, 1.5, 1.0])) == 0


def test_pipeline_c1_chain(gpu_ctx):
    """Config-1 chain through the stage-list API: Demosaic{amaze} -> CST{Srgb}."""
    cfg = im.parse_config('[[pipeline_modules]]\nname = "Demosaic"\nalgorithm = "amaze"\n'
                          '[[pipeline_modules]]\nname = "cst"\ntarget_color_space = "Srgb"\n')
    assert cfg.pipeline_modules == ["Demosaic", "CST"]
    raw, img, plane = _ingest(gpu_ctx, 1024, 1024)
    cal = _calibration()
    img.set_calibration_matrix(cal.ravel())
    im.run_pixel_pipeline(img, cfg)
    want = oracle.camera_cst(oracle.amaze(plane), cal.ravel(), oracle.CS_SRGB)
    got = img.rgb()
    assert np.abs(got - want).max() <= 2e-6
    assert img.color_space == im.CS_SRGB


def test_stage_errors(gpu_ctx):
    img = im.Image(gpu_ctx)
    with pytest.raises(im.ImopsError, match="demosaic"):
        img.cst("Srgb")
    img.set_rgb(np.zeros((4, 4, 3), np.float32), im.CS_NONE)
    with pytest.raises(im.ImopsError, match="calibration matrix"):
        img.cst("Srgb")
    cfg = im.parse_config('[[pipeline_modules]]\nname = "LCH"\n')
    with pytest.raises(im.ImopsError, match="colour space"):
        im.run_pixel_pipeline(img, cfg)
    cfg = im.parse_config('[[pipeline_modules]]\nname = "Demosaic"\nalgorithm = "fast"\n')
    with pytest.raises(im.ImopsError, match="no raw CFA data"):
        im.run_pixel_pipeline(img, cfg)


def test_image_outliving_its_context():
    # freeing an image after its context is destroyed only deletes the handle
    ctx = im.Context(0)
    img = im.Image(ctx).set_rgb(np.ones((8, 8, 3), np.float32), im.CS_NONE)
    ctx.close()
    with pytest.raises(im.ImopsError):
        img.set_rgb(np.ones((8, 8, 3), np.float32), im.CS_NONE)
    img.close()
