"""GPU parity of the film extensions against the oracle, through the C ABI.

The 81-bin spectral grid (380..780 nm / 5 nm), the seeded synthetic stock with
5-knot characteristic curves, Philox4x32-10 grain with a constant density sigma
and the halation radius / strength mapping, in the FP64 form (bar: max abs error
<= 1e-5 on the scanned ACEScg, as for the reference stocks) and in the mixed
FP32 form (bar: PSNR >= 80 dB against the FP64 oracle).  Also through the stage
list's Film fields.
"""
import numpy as np
import pytest

import imops_amd as im
import oracle

pytestmark = pytest.mark.gpu

TOL = 1e-5
C2_FILM = dict(stock="Synthetic", stock_seed=7, spectral_bins=81, grain="philox", grain_sigma=0.01,
               halation_radius=40.0, halation_strength=0.15)


def _scene(w, h, seed):
    g = oracle.absolute_gray(0.185, 200)
    return (im.synth_scene_rgb(w, h, seed) * np.float32(g / 0.185)).astype(np.float32)


def _gpu(ctx, rgb, **kw):
    img = im.Image(ctx).set_rgb(rgb, im.CS_ACESCG)
    img.film(**kw)
    return img.rgb()


@pytest.mark.parametrize("stock", ["ColorNeg200", "Portra400", "TriX400", "EktachromeE100"])
@pytest.mark.parametrize("output", ["NegativeLinear", "PositiveLinear"])
def test_reference_stocks_on_the_81_bin_grid(gpu_ctx, stock, output):
    print(f"film 81 bins {stock} {output}", flush=True)
    rgb = _scene(120, 90, 3)
    got = _gpu(gpu_ctx, rgb, stock=stock, output=output, seed=4, spectral_bins=81)
    want = oracle.film(rgb, stock=stock, output=output, seed=4, spectral_bins=81)
    err = float(np.abs(got - want).max())
    assert err <= TOL, f"max abs err {err}"


@pytest.mark.parametrize("bins", [16, 81])
@pytest.mark.parametrize("stock_seed", [0, 7, 2 ** 40 + 3])
def test_synthetic_stock(gpu_ctx, bins, stock_seed):
    print(f"film synthetic seed {stock_seed} bins {bins}", flush=True)
    rgb = _scene(128, 96, 5)
    for output in ("NegativeLinear", "PositiveLinear"):
        got = _gpu(gpu_ctx, rgb, stock="Synthetic", stock_seed=stock_seed, spectral_bins=bins, output=output, seed=3)
        want = oracle.film(rgb, stock="Synthetic", stock_seed=stock_seed, spectral_bins=bins, output=output, seed=3)
        err = float(np.abs(got - want).max())
        assert err <= TOL, f"{output}: max abs err {err}"


@pytest.mark.parametrize("w,h", [(160, 120), (97, 61)])
def test_c2_film_configuration(gpu_ctx, w, h):
    """The benchmark's film: synthetic stock, 81 bins, Philox grain sigma 0.01, halation 40 px / 0.15."""
    rgb = _scene(w, h, 8)
    got = _gpu(gpu_ctx, rgb, seed=1, **C2_FILM)
    want = oracle.film(rgb, seed=1, **C2_FILM)
    err = float(np.abs(got - want).max())
    assert err <= TOL, f"max abs err {err}"


def test_grain_and_halation_options_each(gpu_ctx):
    rgb = _scene(112, 80, 9)
    for kw in (dict(grain="philox"), dict(grain_sigma=0.02), dict(halation_radius=25.0),
               dict(halation_strength=0.3), dict(halation_radius=12.0, halation_strength=0.0)):
        got = _gpu(gpu_ctx, rgb, stock="ColorNeg200", seed=6, **kw)
        want = oracle.film(rgb, stock="ColorNeg200", seed=6, **kw)
        err = float(np.abs(got - want).max())
        assert err <= TOL, f"{kw}: max abs err {err}"


def test_mixed_precision_81_bins_vs_fp64_oracle(gpu_ctx):
    rgb = _scene(160, 120, 2)
    got = _gpu(gpu_ctx, rgb, seed=1, precision="mixed", **C2_FILM)
    want = oracle.film(rgb, seed=1, fp64=True, **C2_FILM)
    scale = float(np.abs(want).max())
    psnr = 10 * np.log10(scale ** 2 / max(float(np.mean((got - want) ** 2)), 1e-30))
    print(f"mixed 81-bin film PSNR {psnr:.1f} dB", flush=True)
    assert psnr >= 80.0, psnr


def test_film_extension_fields_through_the_stage_list(gpu_ctx):
    rgb = _scene(96, 64, 4)
    cfg = im.parse_config("""
[[pipeline_modules]]
name = "Film"
stock = "Synthetic"
stock_seed = 7
spectral_bins = 81
grain = "philox"
grain_sigma = 0.01
halation_radius = 40
halation_strength = 0.15
seed = 1
""")
    img = im.Image(gpu_ctx).set_rgb(rgb, im.CS_ACESCG)
    im.run_pixel_pipeline(img, cfg)
    want = oracle.film(rgb, seed=1, **C2_FILM)
    assert float(np.abs(img.rgb() - want).max()) <= TOL
    for bad in ('spectral_bins = 40', 'grain = "mt19937"', 'grain_sigma = -1', 'stock_seed = 1.5'):
        with pytest.raises(im.ImopsError):
            im.parse_config(f'[[pipeline_modules]]\nname = "Film"\n{bad}\n')
