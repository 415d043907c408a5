"""GPU parity of the film simulation against the oracle, through the C ABI.

The spectral exposure walk, blurs, grain noise (SplitMix64), DIR/adjacency
algebra and the scan matrices run in the reference's operation order; the
only per-pixel transcendentals evaluated on the device are powf (H&D
reduction), expf (DIR), log10 (capture LUT, f64) and 10^-d as exp(-d ln 10) (scan, f64),
whose device and glibc results may differ by a few f64 ulps.  Bar: max abs error
<= 1e-5 on the scanned ACEScg (values ~0..1), and the grain noise is checked
bit-exact through a flat-field run.
"""
import numpy as np
import pytest

import imops_amd as im
import oracle

pytestmark = pytest.mark.gpu

TOL = 1e-5


def _scene(w, h, seed):
    g = oracle.absolute_gray(0.185, 200)
    return (im.synth_scene_rgb(w, h, seed) * np.float32(g / 0.185)).astype(np.float32)


def _gpu_film(ctx, rgb, **kw):
    img = im.Image(ctx).set_rgb(rgb, im.CS_ACESCG)
    img.film(**kw)
    assert img.color_space == im.CS_ACESCG
    return img.rgb()


@pytest.mark.parametrize("w,h,seed", [(200, 150, 1), (64, 48, 7), (257, 129, 3)])
def test_film_negative_colorneg200(gpu_ctx, w, h, seed):
    print(f"film negative {w}x{h}", flush=True)
    rgb = _scene(w, h, seed)
    got = _gpu_film(gpu_ctx, rgb, seed=seed)
    want = oracle.film(rgb, seed=seed)
    err = np.abs(got - want).max()
    assert err <= TOL, f"max abs err {err}"


STOCKS_ALL = ["BwStub", "ColorNeg200", "Portra400", "Ektar100", "FujiPro400H", "EktachromeE100", "TriX400"]


@pytest.mark.parametrize("stock", STOCKS_ALL)
@pytest.mark.parametrize("output", ["NegativeLinear", "PositiveLinear"])
def test_film_every_stock(gpu_ctx, stock, output):
    """Every stock of film/stock/*.rs (1 to 7 emulsions, masks, reversal, B&W) against the oracle."""
    print(f"film {stock} {output}", flush=True)
    rgb = _scene(96, 72, 11)
    got = _gpu_film(gpu_ctx, rgb, stock=stock, output=output, seed=2)
    want = oracle.film(rgb, stock=stock, output=output, seed=2)
    err = np.abs(got - want).max()
    assert err <= TOL, f"max abs err {err}"


def test_film_positive_linear(gpu_ctx):
    rgb = _scene(160, 120, 2)
    got = _gpu_film(gpu_ctx, rgb, output="PositiveLinear", seed=5)
    want = oracle.film(rgb, output="PositiveLinear", seed=5)
    assert np.abs(got - want).max() <= TOL


def test_film_formats_and_shutter(gpu_ctx):
    rgb = _scene(128, 96, 4)
    for fmt in ("Film6x6", "Film4x5"):
        got = _gpu_film(gpu_ctx, rgb, film_format=fmt)
        want = oracle.film(rgb, film_format=fmt)
        assert np.abs(got - want).max() <= TOL, fmt
    # long exposure: reciprocity failure (radiance.rs:43-49)
    img = im.Image(gpu_ctx).set_rgb(rgb, im.CS_ACESCG).set_shutter_seconds(4.0)
    img.film()
    want = oracle.film(rgb, shutter=4.0)
    assert np.abs(img.rgb() - want).max() <= TOL


def test_film_flat_field_midgray_maps_to_middle_gray(gpu_ctx):
    # PositiveLinear calibrates the mid-gray negative to 0.185 in every channel (invert.rs)
    g = oracle.absolute_gray(0.185, 200)
    rgb = np.full((96, 96, 3), g, np.float32)
    got = _gpu_film(gpu_ctx, rgb, output="PositiveLinear", seed=0)
    assert np.allclose(got.reshape(-1, 3).mean(0), 0.185, atol=2e-3)


def test_film_requires_acescg(gpu_ctx):
    img = im.Image(gpu_ctx).set_rgb(np.ones((8, 8, 3), np.float32), im.CS_SRGB)
    with pytest.raises(im.ImopsError, match="color space"):
        img.film()


def test_c2_chain_small(gpu_ctx):
    """Demosaic{amaze} -> BM3D -> CST{AcesCg} -> Exp -> Film -> CST{Srgb} through the stage list."""
    w, h = 256, 192
    raw = im.synth_raw_u16(w, h, "RGGB", 11)
    cal = np.array([[1.1, -0.1, 0.05], [-0.08, 0.95, 0.1], [0.02, -0.12, 1.2]], np.float32)
    ev = float(np.log2(800.0))
    cfg = im.parse_config(f"""
[[pipeline_modules]]
name = "Demosaic"
algorithm = "amaze"
[[pipeline_modules]]
name = "BM3D"
intensity = 0.1
[[pipeline_modules]]
name = "CST"
target_color_space = "AcesCg"
[[pipeline_modules]]
name = "Exp"
ev = {ev}
[[pipeline_modules]]
name = "Film"
stock = "ColorNeg200"
seed = 3
output = "PositiveLinear"
[[pipeline_modules]]
name = "CST"
target_color_space = "Srgb"
""")
    img = im.Image(gpu_ctx).set_raw_u16(raw, "RGGB", 600.0, 16383.0).set_calibration_matrix(cal.ravel())
    im.run_pixel_pipeline(img, cfg)
    got = img.rgb()

    plane = oracle.normalize_raw(raw, 600.0, 16383.0)
    x = oracle.amaze(plane)
    x, _ = oracle.bm3d(x, oracle.bm3d_params(0.1))
    x = oracle.camera_cst(x, cal.ravel(), oracle.CS_ACESCG)
    x = oracle.exp(x, np.float32(ev))
    x = oracle.film(x, seed=3, output="PositiveLinear")
    x = oracle.cst(x, oracle.CS_ACESCG, oracle.CS_SRGB)
    assert np.abs(got - x).max() <= 2e-5


# The mixed form's film (precision = "mixed"): FP32 per-pixel expose, capture LUT
# and scan instead of the reference's FP64.  Bar: max abs error against the
# oracle below 1e-4 of the output range (single-precision rounding of the 16-bin
# spectral sums, log10 and 10^-d, amplified where the positive inversion takes
# 1/T of a dense negative) and PSNR >= 80 dB, well above the form's 50 dB bar.
@pytest.mark.parametrize("stock,output", [("ColorNeg200", "NegativeLinear"), ("Portra400", "PositiveLinear"),
                                          ("TriX400", "NegativeLinear")])
def test_film_mixed_precision_matches_fp64_oracle(gpu_ctx, stock, output):
    print(f"film mixed {stock} {output}", flush=True)
    rgb = _scene(160, 120, 5)
    cfg = im.parse_config(f"""
[[pipeline_modules]]
name = "Film"
stock = "{stock}"
seed = 2
output = "{output}"
precision = "mixed"
""")
    img = im.Image(gpu_ctx).set_rgb(rgb, im.CS_ACESCG)
    im.run_pixel_pipeline(img, cfg)
    got = img.rgb()
    want = oracle.film(rgb, stock=stock, seed=2, output=output)
    scale = float(np.abs(want).max())
    err = float(np.abs(got - want).max())
    assert err <= 1e-4 * max(scale, 1.0), f"max abs err {err} (output range {scale})"
    psnr = 10 * np.log10(scale ** 2 / max(float(np.mean((got - want) ** 2)), 1e-30))
    assert psnr >= 80.0, psnr


def test_film_precision_rejects_unknown(gpu_ctx):
    with pytest.raises(Exception):
        im.parse_config('[[pipeline_modules]]\nname = "Film"\nprecision = "fp16"\n')


def test_film_seed_beyond_2_pow_53_through_the_stage_list(gpu_ctx):
    """The Film seed is a u64 (film/mod.rs:58): seeds 2^53 and 2^53 + 1 (equal as doubles)
    must give different grain, each the oracle's, when they come through the stage list."""
    rgb = _scene(96, 64, 5)
    outs = {}
    for seed in (2 ** 53, 2 ** 53 + 1, 2 ** 64 - 1):
        cfg = im.parse_config('{"pipeline_modules": [{"name": "Film", "stock": "ColorNeg200", '
                              f'"seed": {seed}}}]}}')
        img = im.Image(gpu_ctx).set_rgb(rgb, im.CS_ACESCG)
        im.run_pixel_pipeline(img, cfg)
        outs[seed] = img.rgb()
        err = np.abs(outs[seed] - oracle.film(rgb, seed=seed)).max()
        assert err <= TOL, f"seed {seed}: max abs err {err}"
    assert not np.array_equal(outs[2 ** 53], outs[2 ** 53 + 1])
