"""GPU parity of BM3D with the Haar group transform (group_transform = "haar",
an extension beside the reference's Walsh-Hadamard), through the C ABI.

Bars: the plain FP32 form bit-exact against the FP32 oracle (same transform,
same operation order); groups of two bit-exact against the Walsh-Hadamard
output; "mixed-match" bit-exact against the FP32 oracle when no patch is
selected approximately; the mixed form >= 50 dB against the FP64 oracle.
"""
import numpy as np
import pytest

import imops_amd as im
import oracle

pytestmark = pytest.mark.gpu


def _scene(w, h, seed, sigma=0.03):
    rng = np.random.default_rng(seed)
    return np.clip(im.synth_scene_rgb(w, h, seed) + rng.normal(0, sigma, (h, w, 3)), 0, 1).astype(np.float32)


def _gpu(ctx, rgb, intensity, **explicit):
    img = im.Image(ctx).set_rgb(rgb, im.CS_NONE)
    ctx.set_timing(True)
    img.bm3d(intensity, **explicit)
    approx, _ = ctx.timing("bm3d_approx_refs")
    ctx.set_timing(False)
    return img.rgb(), int(approx)


def _psnr(a, b):
    mse = float(np.mean((a.astype(np.float64) - b.astype(np.float64)) ** 2))
    return float("inf") if mse == 0 else 10 * np.log10(1.0 / mse)


def _bitexact(got, want):
    bad = np.count_nonzero(got.view(np.uint32) != want.view(np.uint32))
    assert bad == 0, f"{bad} mismatching values, max abs {np.abs(got - want).max()}"


CASES = [
    ("ref-params", 96, 80, 0.1, {}, {}),
    ("ref-params-strong", 131, 67, 0.3, {}, {}),
    ("config-geometry", 120, 96, 0.06375, dict(sigma=5.1, lambda_=2.7, stride=3, search_window=39),
     dict(sigma=5.1, stride=3, window=39, **{"lambda": 2.7})),
    ("group-4", 100, 90, 0.2, dict(search_window=39, max_group=4), dict(window=39, max_group=4)),
    ("group-8-cap-40", 100, 90, 0.2, dict(search_window=39, max_group=8, max_candidates=40),
     dict(window=39, max_group=8, max_candidates=40)),
]


@pytest.mark.parametrize("name,w,h,intensity,explicit,oparams", CASES, ids=[c[0] for c in CASES])
def test_bm3d_haar_fp32_bitexact(gpu_ctx, name, w, h, intensity, explicit, oparams):
    print(f"bm3d haar fp32 {name} {w}x{h}", flush=True)
    rgb = _scene(w, h, seed=w + 3 * h)
    want, _ = oracle.bm3d(rgb, oracle.bm3d_params(intensity, group_transform="haar", **oparams))
    got, _ = _gpu(gpu_ctx, rgb, intensity, group_transform="haar", **explicit)
    _bitexact(got, want)


def test_bm3d_haar_groups_of_two_equal_hadamard(gpu_ctx):
    rgb = _scene(96, 72, seed=9)
    a, _ = _gpu(gpu_ctx, rgb, 0.3, group_transform="haar", max_group=2)
    b, _ = _gpu(gpu_ctx, rgb, 0.3, max_group=2)
    _bitexact(a, b)
    c, _ = _gpu(gpu_ctx, rgb, 0.3, group_transform="haar")
    d, _ = _gpu(gpu_ctx, rgb, 0.3)
    assert not np.array_equal(c, d)


@pytest.mark.parametrize("name,w,h,intensity,explicit,oparams", CASES, ids=[c[0] for c in CASES])
def test_bm3d_haar_mixed(gpu_ctx, name, w, h, intensity, explicit, oparams):
    print(f"bm3d haar mixed {name} {w}x{h}", flush=True)
    rgb = _scene(w, h, seed=w + 3 * h)
    p = oracle.bm3d_params(intensity, group_transform="haar", **oparams)
    want, _ = oracle.bm3d(rgb, p)
    want64, _ = oracle.bm3d(rgb, p, fp64=True)
    got, approx = _gpu(gpu_ctx, rgb, intensity, group_transform="haar", precision="mixed-match", **explicit)
    if approx == 0:
        _bitexact(got, want)
    got, _ = _gpu(gpu_ctx, rgb, intensity, group_transform="haar", precision="mixed", **explicit)
    p64 = _psnr(got, want64)
    print(f"  mixed vs FP64 oracle {p64:.1f} dB", flush=True)
    assert p64 >= 50.0


def test_bm3d_haar_through_stage_list(gpu_ctx):
    rgb = _scene(120, 96, seed=5)
    cfg = im.parse_config('[[pipeline_modules]]\nname = "BM3D"\nintensity = 0.2\ngroup_transform = "haar"\n')
    img = im.Image(gpu_ctx).set_rgb(rgb, im.CS_NONE)
    im.run_pixel_pipeline(img, cfg)
    want, _ = oracle.bm3d(rgb, oracle.bm3d_params(0.2, group_transform="haar"))
    _bitexact(img.rgb(), want)
