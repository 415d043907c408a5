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


def _opcodes(k, cx, cy):
    """A DNG OpcodeList3 with one FixVignetteRadial entry (after a skipped opcode)."""
    import struct
    body = struct.pack(">7d", *k, cx, cy)
    skip = struct.pack(">IIII", 1, 1, 0, 4) + b"\0" * 4
    return struct.pack(">I", 2) + skip + struct.pack(">IIII", 3, 1, 0, len(body)) + body


def test_burst_graphs_keep_per_frame_exposure_and_vignette_metadata(gpu_ctx):
    """Frames whose stage list reads metadata (BaselineExposureCompensation: baseline
    EV, f-number, ISO, shutter; Vignette: OpcodeList3).  Frames with the same metadata
    share a graph; every replayed frame must get its own gains and vignette, bit for
    bit as when run alone."""
    cfg = im.parse_config("""
[[pipeline_modules]]
name = "Demosaic"
algorithm = "markesteijn"
[[pipeline_modules]]
name = "CFACoeffs"
[[pipeline_modules]]
name = "Vignette"
strength = 0.8
[[pipeline_modules]]
name = "BaselineExposureCompensation"
[[pipeline_modules]]
name = "CST"
target_color_space = "Srgb"
""")
    ops_a = _opcodes((-0.4, 0.15, 0.02, 0.0, -0.01), 0.47, 0.55)
    ops_b = _opcodes((-0.2, 0.05, 0.0, 0.0, 0.0), 0.5, 0.5)
    # (seed, ev, f-number, iso, shutter, opcodes): groups of 3, 2 and 1 frames
    meta = [(1, 0.35, 5.6, 200.0, 1 / 125, ops_a), (2, 0.35, 5.6, 200.0, 1 / 125, ops_a),
            (3, -0.5, 2.8, 800.0, 1 / 30, ops_b), (4, 0.35, 5.6, 200.0, 1 / 125, ops_a),
            (5, -0.5, 2.8, 800.0, 1 / 30, ops_b), (6, 0.0, 8.0, 100.0, 1 / 500, None)]

    def frame(seed, ev, fn, iso, sh, ops):
        f = _frame(gpu_ctx, 288, 192, "RGGB", seed).set_wb_coeffs([2.0, 1.0, 1.5, 1.0])
        f.set_shutter_seconds(sh).set_exposure_metadata(ev, fn, iso)
        if ops is not None:
            f.set_opcode_list3(ops)
        return f

    eager = []
    for m in meta:
        f = frame(*m)
        im.run_pixel_pipeline(f, cfg)
        eager.append(f.rgb())
    frames = [frame(*m) for m in meta]
    im.run_burst(frames, cfg, graphs=True)
    for i, (f, want) in enumerate(zip(frames, eager)):
        got = f.rgb()
        bad = np.count_nonzero(got.view(np.uint32) != want.view(np.uint32))
        print(f"metadata burst frame {i}: {bad} differing values", flush=True)
        assert bad == 0


@pytest.mark.parametrize("algorithm,factor", [("fast", 2), ("superfast", 4)])
def test_burst_graphs_with_a_binning_preview_demosaic(gpu_ctx, algorithm, factor):
    """A preview demosaic bins the frame (w/f x h/f): every replayed frame must come
    back with the binned shape and the same bits as when run alone."""
    cfg = im.parse_config(f"""
[[pipeline_modules]]
name = "Demosaic"
algorithm = "{algorithm}"
[[pipeline_modules]]
name = "CST"
target_color_space = "Srgb"
""")
    specs = [(288, 192, 31), (288, 192, 32), (192, 160, 33), (288, 192, 34), (192, 160, 35)]
    eager = []
    for w, h, seed in specs:
        f = _frame(gpu_ctx, w, h, "RGGB", seed)
        im.run_pixel_pipeline(f, cfg)
        eager.append(f.rgb())
    frames = [_frame(gpu_ctx, w, h, "RGGB", seed) for w, h, seed in specs]
    im.run_burst(frames, cfg, graphs=True)
    for (w, h, _), f, want in zip(specs, frames, eager):
        assert (f.width, f.height) == (w // factor, h // factor)
        got = f.rgb()
        assert got.shape == want.shape
        assert np.array_equal(got.view(np.uint32), want.view(np.uint32))
