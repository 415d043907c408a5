"""GPU parity of the mixed-precision BM3D forms.

"mixed-match": FP16 matrix-core block matching, FP32 filter.  The matcher
carries a rigorous error interval for every FP16 distance and re-evaluates, in
the reference's FP32 order, every candidate whose interval reaches the
acceptance threshold or the group boundary; such patches get exactly the FP32
matcher's group.  Where more than 48 candidates are indistinguishable at FP16
precision (near-flat guides), the group is chosen from the approximate top 15
instead, re-ranked in FP32 (the "approximate" selection).

"mixed": the same matcher plus the collaborative filter's 2-D DCT and
Walsh-Hadamard transforms as FP16 matrix-core products with FP32 accumulation
(thresholding, Wiener shrinkage, weights and aggregation stay FP32 / fixed point).

Bars: "mixed-match" bit-exact with the FP32 oracle when the library reports no
approximate selection; both forms PSNR >= 50 dB against the FP64 oracle (the
mixed form's target).  The cases cover the reference geometry (stride 6,
19x19), the 39x39 / stride 3 geometry where the 1024-candidate cap binds, small
caps and group limits, thresholds tight enough that many intervals straddle
them, flat and saturated areas, and the reference's golden checksum image.
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


def _mixed(ctx, rgb, intensity, precision="mixed-match", **explicit):
    """Mixed-form output and the number of approximately selected patches."""
    img = im.Image(ctx).set_rgb(rgb, im.CS_NONE)
    ctx.set_timing(True)
    img.bm3d(intensity, precision=precision, **explicit)
    approx, _ = ctx.timing("bm3d_approx_refs")
    ctx.set_timing(False)
    return img.rgb(), int(approx)


def _psnr(a, b):
    mse = float(np.mean((a.astype(np.float64) - b.astype(np.float64)) ** 2))
    return float("inf") if mse == 0 else 10 * np.log10(1.0 / mse)


def _check(got, approx, rgb, params, exact=True):
    want, _ = oracle.bm3d(rgb, params)
    want64, _ = oracle.bm3d(rgb, params, fp64=True)
    p64 = _psnr(got, want64)
    print(f"approximate patches {approx}, PSNR vs FP64 oracle {p64:.1f} dB, vs FP32 oracle {_psnr(got, want):.1f} dB",
          flush=True)
    assert p64 >= 50.0
    if exact and approx == 0:
        bad = np.count_nonzero(got.view(np.uint32) != want.view(np.uint32))
        assert bad == 0, f"{bad} mismatching values, max abs {np.abs(got - want).max()}"


CASES = [
    ("ref-params", 131, 97, 0.1, {}),
    ("ref-params-strong", 96, 80, 0.8, {}),
    ("config-geometry", 200, 150, 0.06375, dict(sigma=5.1, lambda_=2.7, stride=3, search_window=39)),
    ("cap-40", 100, 90, 0.2, dict(search_window=39, max_candidates=40)),
    ("cap-7-group-8", 100, 90, 0.2, dict(search_window=39, max_candidates=7, max_group=8)),
    ("group-1", 100, 90, 0.2, dict(search_window=39, max_group=1)),
    ("tight-threshold", 120, 100, 0.1, dict(max_dist=60.0)),
    ("tight-threshold-39", 120, 100, 0.1, dict(max_dist=45.0, stride=3, search_window=39)),
    ("stride-1-window-9", 70, 50, 0.3, dict(stride=1, search_window=9)),
    ("thin", 300, 12, 0.1, dict(stride=2, search_window=21)),
]


def _case_params(intensity, explicit):
    p = oracle.bm3d_params(intensity, **{k: v for k, v in explicit.items() if k not in ("lambda_", "search_window")},
                           **({"window": explicit["search_window"]} if "search_window" in explicit else {}))
    if "lambda_" in explicit:
        p["lambda"] = explicit["lambda_"]
    return p


@pytest.mark.parametrize("name,w,h,intensity,explicit", CASES, ids=[c[0] for c in CASES])
def test_bm3d_mixed_match_equals_fp32_oracle(gpu_ctx, name, w, h, intensity, explicit):
    print(f"bm3d mixed-match {name} {w}x{h}", flush=True)
    rgb = _noisy_scene(w, h, seed=w * 7 + h)
    _check(*_mixed(gpu_ctx, rgb, intensity, **explicit), rgb, _case_params(intensity, explicit))


@pytest.mark.parametrize("name,w,h,intensity,explicit", CASES, ids=[c[0] for c in CASES])
def test_bm3d_mixed_filter_psnr(gpu_ctx, name, w, h, intensity, explicit):
    print(f"bm3d mixed {name} {w}x{h}", flush=True)
    rgb = _noisy_scene(w, h, seed=w * 7 + h)
    _check(*_mixed(gpu_ctx, rgb, intensity, precision="mixed", **explicit), rgb, _case_params(intensity, explicit),
           exact=False)


@pytest.mark.parametrize("precision", ["mixed-match", "mixed"])
def test_bm3d_mixed_flat_and_saturated(gpu_ctx, precision):
    # large flat areas make many equal distances (ties resolved by scan order)
    rgb = np.zeros((90, 120, 3), np.float32)
    rgb[:, :60] = 0.5
    rgb[30:60, 40:100] = 1.0
    rgb[::7, ::5] = 0.25
    _check(*_mixed(gpu_ctx, rgb, 0.2, precision=precision), rgb, oracle.bm3d_params(0.2),
           exact=precision == "mixed-match")


@pytest.mark.parametrize("precision", ["mixed-match", "mixed"])
def test_bm3d_mixed_golden_checksums(gpu_ctx, precision):
    """bm3d.rs:791-808 image through the mixed forms ("mixed-match": same bits as the FP32 form)."""
    img = oracle.bm3d_test_image(1024, 1024)
    _check(*_mixed(gpu_ctx, img, 0.1, precision=precision), img, oracle.bm3d_params(0.1),
           exact=precision == "mixed-match")


@pytest.mark.parametrize("precision", ["mixed-match", "mixed"])
def test_bm3d_mixed_through_stage_list(gpu_ctx, precision):
    rgb = _noisy_scene(100, 64, seed=9)
    cfg = im.parse_config(f'[[pipeline_modules]]\nname = "BM3D"\nintensity = 0.15\nprecision = "{precision}"\n')
    img = im.Image(gpu_ctx).set_rgb(rgb, im.CS_NONE)
    gpu_ctx.set_timing(True)
    im.run_pixel_pipeline(img, cfg)
    approx, _ = gpu_ctx.timing("bm3d_approx_refs")
    gpu_ctx.set_timing(False)
    _check(img.rgb(), int(approx), rgb, oracle.bm3d_params(0.15), exact=precision == "mixed-match")


def test_bm3d_mixed_match_non_finite_pixels(gpu_ctx):
    """NaN and +-Inf pixels would turn FP16 distances into NaN, which the
    matrix-core sweeps read as "outside the window" (such a pair would count as
    accepted); the blocks whose staged region holds a value that is not finite
    or past FP16's range go to the plain matcher, so "mixed-match" keeps the FP32
    form's bits.  (Finite values that fit FP16 but dwarf the rest, such as RGB
    300, widen the block's error bounds until its selections are approximate,
    which the library reports; that case is not exact by design.)"""
    rgb = _noisy_scene(120, 96, seed=21)
    rng = np.random.default_rng(22)
    for v in (np.nan, np.inf, -np.inf):
        ys, xs = rng.integers(0, 96, 4), rng.integers(0, 120, 4)
        rgb[ys, xs, rng.integers(0, 3, 4)] = v
    p = oracle.bm3d_params(0.1, stride=3, window=39)
    got, approx = _mixed(gpu_ctx, rgb, 0.1, stride=3, search_window=39)
    want, _ = oracle.bm3d(rgb, p)
    assert approx == 0
    bad = np.count_nonzero(got.view(np.uint32) != want.view(np.uint32))
    assert bad == 0, f"{bad} mismatching values"
