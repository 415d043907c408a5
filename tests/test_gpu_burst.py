"""Bursts of uneven frames through one stage list (imops_pipeline_run_burst).

The graph form (each shape after its first frame replays a HIP graph captured on
a staging image) must give every frame the same bits as running the stage list
on that frame alone.  The bursts mix sizes (and, with the 3x3 same-colour
demosaic, Bayer and X-Trans patterns), repeat shapes so every graph replays
several times with the library's own eager work in between, and run the whole
burst twice.
"""
import numpy as np
import pytest

import imops_amd as im

pytestmark = pytest.mark.gpu

XTRANS = "GGRGGB" "GGBGGR" "BRGRBG" "GGBGGR" "GGRGGB" "RBGBRG"
CAL = np.array([[1.1, -0.1, 0.05], [-0.08, 0.95, 0.1], [0.02, -0.12, 1.2]], np.float32)


def _stages(demosaic, precision):
    return f"""
[[pipeline_modules]]
name = "Demosaic"
algorithm = "{demosaic}"

[[pipeline_modules]]
name = "BM3D"
intensity = 0.1
precision = "{precision}"

[[pipeline_modules]]
name = "CST"
target_color_space = "AcesCg"

[[pipeline_modules]]
name = "Exp"
ev = 9.64

[[pipeline_modules]]
name = "Film"
stock = "ColorNeg200"
seed = 3

[[pipeline_modules]]
name = "CST"
target_color_space = "Srgb"
"""


def _frame(ctx, w, h, cfa, seed):
    raw = im.synth_raw_u16(w, h, cfa, seed)
    return im.Image(ctx).set_raw_u16(raw, cfa, 600.0, 16383.0).set_calibration_matrix(CAL.ravel())


def _run_burst(ctx, specs, demosaic, precision):
    cfg = im.parse_config(_stages(demosaic, precision))
    eager = []
    for w, h, cfa, seed in specs:
        f = _frame(ctx, w, h, cfa, seed)
        im.run_pixel_pipeline(f, cfg)
        eager.append(f.rgb())
    for rep in range(2):
        frames = [_frame(ctx, w, h, cfa, seed) for w, h, cfa, seed in specs]
        im.run_burst(frames, cfg, graphs=True)
        for i, (f, want) in enumerate(zip(frames, eager)):
            got = f.rgb()
            bad = np.count_nonzero(got.view(np.uint32) != want.view(np.uint32))
            print(f"burst rep {rep} frame {i} {specs[i][:3]}: {bad} differing values", flush=True)
            assert bad == 0


@pytest.mark.parametrize("precision", ["fp32", "mixed"])
def test_burst_graphs_bayer_uneven_sizes(gpu_ctx, precision):
    specs = [(288, 192, "RGGB", 1), (192, 160, "RGGB", 2), (288, 192, "RGGB", 3), (384, 256, "RGGB", 4),
             (192, 160, "RGGB", 5), (288, 192, "RGGB", 6)]
    _run_burst(gpu_ctx, specs, "amaze", precision)


def test_burst_graphs_mixed_cfa(gpu_ctx):
    specs = [(312, 208, XTRANS, 11), (288, 192, "RGGB", 12), (312, 208, XTRANS, 13), (288, 192, "RGGB", 14),
             (312, 208, XTRANS, 15)]
    _run_burst(gpu_ctx, specs, "markesteijn", "mixed")


def test_burst_without_graphs_matches(gpu_ctx):
    specs = [(288, 192, "RGGB", 21), (192, 160, "RGGB", 22), (288, 192, "RGGB", 23)]
    cfg = im.parse_config(_stages("amaze", "fp32"))
    frames = [_frame(gpu_ctx, w, h, cfa, seed) for w, h, cfa, seed in specs]
    im.run_burst(frames, cfg, graphs=False)
    for (w, h, cfa, seed), f in zip(specs, frames):
        ref = _frame(gpu_ctx, w, h, cfa, seed)
        im.run_pixel_pipeline(ref, cfg)
        assert np.array_equal(f.rgb().view(np.uint32), ref.rgb().view(np.uint32))
