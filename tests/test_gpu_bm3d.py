"""GPU parity of BM3D (plain FP32) against the oracle, through the C ABI.

Bar: bit-exact.  Matching sums, transforms and the Wiener running sum follow
the reference's f32 operation order and aggregation is exact i64 arithmetic,
so the GPU output must equal the oracle's bit for bit (ties in the match sort
are broken by scan order on both sides).
"""
import numpy as np
import pytest

import imops_amd as im
import oracle

pytestmark = pytest.mark.gpu


def _noisy_scene(w, h, seed, sigma=0.03):
    rng = np.random.default_rng(seed)
    x = im.synth_scene_rgb(w, h, seed) + rng.normal(0, sigma, (h, w, 3)).astype(np.float32)
    return np.clip(x, 0, 1).astype(np.float32)


def _gpu(ctx, rgb, intensity=None, **explicit):
    img = im.Image(ctx).set_rgb(rgb, im.CS_NONE)
    if explicit:
        img.bm3d(intensity or 0.0, **explicit)
    else:
        img.bm3d(intensity)
    return img.rgb()


def _check(got, want):
    bad = np.count_nonzero(got.view(np.uint32) != want.view(np.uint32))
    assert bad == 0, f"{bad} mismatching values, max abs {np.abs(got - want).max()}"


@pytest.mark.parametrize("w,h,intensity", [(96, 80, 0.1), (131, 67, 0.3), (64, 64, 1.0), (40, 30, 0.02)])
def test_bm3d_reference_params(gpu_ctx, w, h, intensity):
    print(f"bm3d {w}x{h} i={intensity}", flush=True)
    rgb = _noisy_scene(w, h, seed=w + h)
    want, _ = oracle.bm3d(rgb, oracle.bm3d_params(intensity))
    _check(_gpu(gpu_ctx, rgb, intensity), want)


def test_bm3d_config_geometry(gpu_ctx):
    # stride 3, 39x39 window, lambda 2.7, sigma 0.02 normalised (5.1 in 0..255 units)
    rgb = _noisy_scene(120, 96, seed=3, sigma=0.02)
    i = 0.02 * 255 / 80
    p = oracle.bm3d_params(i, sigma=0.02 * 255, stride=3, window=39)
    p["lambda"] = 2.7
    want, _ = oracle.bm3d(rgb, p)
    got = _gpu(gpu_ctx, rgb, i, sigma=np.float32(0.02 * 255), lambda_=2.7, stride=3, search_window=39)
    _check(got, want)


def _stripes(w, h, seed):
    rng = np.random.default_rng(seed)
    g = rng.uniform(0.1, 0.9, (8, 3))
    c = rng.uniform(-0.1, 0.1, ((w + 11) // 12, 3))
    y, x = np.arange(h)[:, None], np.arange(w)[None, :]
    return np.clip(g[y % 8] + c[x // 12], 0, 1).astype(np.float32)


@pytest.mark.parametrize("window", [39, 19])
def test_bm3d_exact_copies_eight_rows_apart(gpu_ctx, window):
    """Rows repeat every 8 lines and columns are flat over 12-pixel segments, so a
    patch's exact copies (distance 0, ties broken by scan order) lie 8 rows apart
    in the same few columns.  At 39x39 the matcher's lanes walk 8 blocks of 5
    candidates per window row, so those copies fall to one lane: more of the
    group than the 8 keys a lane keeps, which makes the wave walk the window
    again for that lane's next keys (a CPU model of the selection counts 262 of
    the first 300 interior patches needing it)."""
    rgb = _stripes(120, 100, 7)
    i = 0.02 * 255 / 80
    p = oracle.bm3d_params(i, sigma=0.02 * 255, stride=3, window=window)
    p["lambda"] = 2.7
    want, _ = oracle.bm3d(rgb, p)
    got = _gpu(gpu_ctx, rgb, i, sigma=np.float32(0.02 * 255), lambda_=2.7, stride=3, search_window=window)
    _check(got, want)


def test_bm3d_candidate_cap_and_group_limits(gpu_ctx):
    rgb = _noisy_scene(100, 90, seed=5)
    for cap, grp in [(40, 16), (1024, 4), (7, 8), (1024, 1)]:
        print(f"cap={cap} group={grp}", flush=True)
        p = oracle.bm3d_params(0.2, window=39, max_candidates=cap, max_group=grp)
        want, _ = oracle.bm3d(rgb, p)
        got = _gpu(gpu_ctx, rgb, 0.2, search_window=39, max_candidates=cap, max_group=grp)
        _check(got, want)


def test_bm3d_tiny_image_no_patches(gpu_ctx):
    rgb = _noisy_scene(8, 7, seed=1)
    want, _ = oracle.bm3d(rgb, oracle.bm3d_params(0.1))
    _check(_gpu(gpu_ctx, rgb, 0.1), want)


def test_bm3d_reference_golden_checksums(gpu_ctx):
    """bm3d.rs:791-808: 2048x2048 sinusoids, intensity 0.1, Kahan channel sums.

    The GPU must equal the oracle bit for bit; both land within 0.05 of the
    reference's recorded sums (the residual ~0.02-0.03 is the size of the shift
    that one-ulp differences in the fixture's sinf/cosf produce)."""
    img = oracle.bm3d_test_image(2048, 2048)
    got = _gpu(gpu_ctx, img, 0.1)
    want, _ = oracle.bm3d(img, oracle.bm3d_params(0.1))
    _check(got, want)
    expected = [2100044.776915211, 2101586.5654235613, 2098591.6187395807]
    for c in range(3):
        assert abs(oracle.refined_sum(got, c) - expected[c]) < 0.05


def test_bm3d_through_stage_list(gpu_ctx):
    rgb = _noisy_scene(100, 64, seed=9)
    cfg = im.parse_config('[[pipeline_modules]]\nname = "BM3D"\nintensity = 0.15\n')
    img = im.Image(gpu_ctx).set_rgb(rgb, im.CS_NONE)
    im.run_pixel_pipeline(img, cfg)
    want, _ = oracle.bm3d(rgb, oracle.bm3d_params(0.15))
    _check(img.rgb(), want)


@pytest.mark.parametrize("stride,window", [(6, 41), (10, 41), (17, 41), (1, 41), (2, 1), (4, 7)])
def test_bm3d_staging_layouts(gpu_ctx, stride, window):
    """Strides and windows that give every staging layout of the matcher: four,
    three, two and one reference patches per workgroup sharing a staged window
    ((R - 1) * stride + window + 7 <= 64 columns), a 1x1 window and a 7x7 one."""
    print(f"bm3d staging stride {stride} window {window}", flush=True)
    rgb = _noisy_scene(90, 70, seed=stride * 100 + window)
    p = oracle.bm3d_params(0.2, stride=stride, window=window)
    want, _ = oracle.bm3d(rgb, p)
    _check(_gpu(gpu_ctx, rgb, 0.2, stride=stride, search_window=window), want)


@pytest.mark.parametrize("window", [39, 19])
def test_bm3d_nan_and_inf_pixels(gpu_ctx, window):
    """NaN and +-Inf pixels: the reference's row sums stop at the first row end
    above the threshold (an Inf row stops it at once) while the GPU sums every
    row (an Inf minus an Inf later in the patch makes its full sum NaN); both
    reject such candidates, so the groups, and the fixed-point aggregation that
    turns NaN contributions into zero, agree bit for bit."""
    rgb = _noisy_scene(96, 80, seed=11)
    rng = np.random.default_rng(12)
    for v in (np.nan, np.inf, -np.inf):
        ys, xs = rng.integers(0, 80, 6), rng.integers(0, 96, 6)
        rgb[ys, xs, rng.integers(0, 3, 6)] = v
    p = oracle.bm3d_params(0.1, stride=3, window=window)
    want, _ = oracle.bm3d(rgb, p)
    got = _gpu(gpu_ctx, rgb, 0.1, stride=3, search_window=window)
    bad = np.count_nonzero(got.view(np.uint32) != want.view(np.uint32))
    assert bad == 0, f"{bad} mismatching values"
