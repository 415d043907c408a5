"""Launch-shape autotuning (imops_ctx_set_autotune): every candidate computes the
same result, so BM3D with the tuner on must give the same bits as with it off,
record one choice per tunable kernel and shape, and reuse it on later calls."""
import numpy as np
import pytest

import imops_amd as im

pytestmark = pytest.mark.gpu


def _noisy_scene(w, h, seed, sigma=0.03):
    rng = np.random.default_rng(seed)
    x = im.synth_scene_rgb(w, h, seed) + rng.normal(0, sigma, (h, w, 3)).astype(np.float32)
    return np.clip(x, 0, 1).astype(np.float32)


def _bm3d(ctx, rgb, precision):
    img = im.Image(ctx).set_rgb(rgb, im.CS_NONE)
    img.bm3d(0.06375, sigma=np.float32(5.1), lambda_=2.7, stride=3, search_window=39, precision=precision)
    return img.rgb()


@pytest.mark.parametrize("precision", ["fp32", "mixed"])
def test_autotuned_bm3d_is_bit_identical(gpu_ctx, precision):
    rgb = _noisy_scene(200, 150, 5)
    want = _bm3d(gpu_ctx, rgb, precision)
    tuned_ctx = im.Context(0)
    try:
        tuned_ctx.set_autotune(True)
        got = _bm3d(tuned_ctx, rgb, precision)
        choices = tuned_ctx.tuning()
        print(f"{precision}: {choices}", flush=True)
        assert np.array_equal(got.view(np.uint32), want.view(np.uint32))
        kinds = {k.split("/")[0] for k in choices}
        assert ("bm3d_filter_mfma" if precision == "mixed" else "bm3d_filter") in kinds
        if precision == "fp32":
            assert "bm3d_match" in kinds
        assert all(v in (0, 2, 4, 5, 7, 8) for v in choices.values())
        again = _bm3d(tuned_ctx, rgb, precision)  # cached choices, no new trials
        assert tuned_ctx.tuning() == choices
        assert np.array_equal(again.view(np.uint32), want.view(np.uint32))
    finally:
        tuned_ctx.close()
