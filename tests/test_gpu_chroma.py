"""GPU parity of the chroma denoise stages against the oracle, through the C ABI.

ChromaDenoise (chroma_bm3d, bm3d.rs:705-759): the DCTs follow the reference's
f32 operation order and the Kaiser aggregation is exact i64 arithmetic, so the
bar is bit-exact; the reference's own 2048x2048 checksums (bm3d.rs:810-827)
are checked on the GPU output as well.
"""
import numpy as np
import pytest

import imops_amd as im
import oracle

pytestmark = pytest.mark.gpu


def _noisy_scene(w, h, seed, sigma=0.04):
    rng = np.random.default_rng(seed)
    x = im.synth_scene_rgb(w, h, seed) + rng.normal(0, sigma, (h, w, 3)).astype(np.float32)
    return np.clip(x, 0, 1).astype(np.float32)


def _check(got, want):
    bad = np.count_nonzero(got.view(np.uint32) != want.view(np.uint32))
    assert bad == 0, f"{bad} mismatching values, max abs {np.abs(got - want).max()}"


# ragged sizes around the 90-pixel tile and the 6-pixel patch step; 8 and 5
# wide frames have no patch at all (only the YCbCr round trip)
@pytest.mark.parametrize("w,h,intensity", [(200, 150, 0.1), (181, 97, 0.5), (96, 90, 1.0), (15, 14, 0.02),
                                           (8, 40, 0.3), (5, 3, 0.3), (91, 181, 0.001)])
def test_chroma_bm3d_matches_oracle(gpu_ctx, w, h, intensity):
    print(f"chroma_bm3d {w}x{h} i={intensity}", flush=True)
    rgb = _noisy_scene(w, h, seed=w * 7 + h)
    got = im.Image(gpu_ctx).set_rgb(rgb, im.CS_NONE).chroma_bm3d(intensity).rgb()
    _check(got, oracle.chroma_bm3d(rgb, intensity))


def test_chroma_bm3d_reference_golden_checksums(gpu_ctx):
    img = oracle.bm3d_test_image(2048, 2048)
    got = im.Image(gpu_ctx).set_rgb(img, im.CS_NONE).chroma_bm3d(0.1).rgb()
    expected = [2099732.10061259, 2101692.5347650475, 2098392.755773447]
    for c in range(3):
        assert abs(oracle.refined_sum(got, c) - expected[c]) < 3e-3
    _check(got, oracle.chroma_bm3d(img, 0.1))


def test_chroma_denoise_stage_in_pipeline(gpu_ctx):
    rgb = _noisy_scene(160, 120, 3)
    cfg = im.parse_config('[[pipeline_modules]]\nname = "chromadenoise"\nintensity = 0.4\n')
    img = im.Image(gpu_ctx).set_rgb(rgb, im.CS_NONE)
    im.run_pixel_pipeline(img, cfg)
    _check(img.rgb(), oracle.chroma_bm3d(rgb, 0.4))


# LumaGuidedChromaDenoise (chroma_denoise.rs:14-152): the box filter's running
# sums are walked in the reference's order, one lane per line, so the bar is
# bit-exact.  Sizes cross the 64-line / 64-position chunks; radii include 0,
# radii past a chunk and past the frame (clamped windows), and 1-pixel frames.
@pytest.mark.parametrize("w,h,radius,eps", [(200, 150, 4, 0.01), (131, 67, 0, 0.01), (64, 300, 8, 0.05),
                                            (97, 65, 70, 0.01), (1, 50, 3, 0.01), (70, 1, 2, 0.01),
                                            (300, 200, 33, 0.1), (129, 129, 64, 0.01)])
def test_luma_chroma_denoise_matches_oracle(gpu_ctx, w, h, radius, eps):
    print(f"luma_chroma_denoise {w}x{h} r={radius} eps={eps}", flush=True)
    rgb = _noisy_scene(w, h, seed=w + 3 * h)
    got = im.Image(gpu_ctx).set_rgb(rgb, im.CS_NONE).chroma_denoise(radius, eps).rgb()
    _check(got, oracle.luma_chroma_denoise(rgb, radius, eps))


def test_luma_guided_stage_in_pipeline(gpu_ctx):
    rgb = _noisy_scene(150, 100, 9)
    cfg = im.parse_config('[[pipeline_modules]]\nname = "LumaGuidedChromaDenoise"\nradius = 5\nepsilon = 0.02\n')
    img = im.Image(gpu_ctx).set_rgb(rgb, im.CS_NONE)
    im.run_pixel_pipeline(img, cfg)
    _check(img.rgb(), oracle.luma_chroma_denoise(rgb, 5, np.float32(0.02)))
