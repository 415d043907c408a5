"""The benchmark workloads' stage lists, run through the C ABI on crops of their
frames, against the oracle.

C2 (the headline bench): Demosaic{amaze} -> BM3D{stride 3, 39x39 window, <= 16
matches, lambda 2.7, sigma 0.02} -> CST{AcesCg} -> Exp{log2 800} -> Film ->
CST{Srgb}, exactly the text bench.py times, on a 768x512 crop of the
6016x4016 RGGB frame.  C3: the same list with Demosaic{markesteijn} on a crop
of the 6240x4160 X-Trans frame.  C5: a burst of crops of the four burst frame
sizes (per-frame read noise 1-8 DN) through graph replays.

Bars: the plain FP32 form <= 2e-5 max abs against the FP32 exact-order oracle
(bit-exact stages plus the device powf / exp / log10 of the sRGB OETF, the
H&D reduction, DIR and the film scan); the mixed form (FP16 matrix-core BM3D,
FP32 film spectral arithmetic) >= 50 dB PSNR against the FP64 oracle.
"""
import numpy as np
import pytest

import bench
import imops_amd as im

pytestmark = pytest.mark.gpu

TOL_FP32 = 2e-5
PSNR_MIXED_DB = 50.0
CW, CH = bench.CROP_W, bench.CROP_H


@pytest.fixture(scope="module")
def cal():
    return bench.calibration_matrix()


def _gpu_chain(ctx, raw, cfa, cal, precision, demosaic):
    img = im.Image(ctx).set_calibration_matrix(cal.ravel()).set_raw_u16(raw, cfa, bench.BLACK, bench.WHITE)
    im.run_pixel_pipeline(img, im.parse_config(bench.stages_demosaic(precision, demosaic)))
    return img.rgb()


def _check(got, raw, cfa, cal, precision, demosaic, what):
    if precision == "fp32":
        want = bench.oracle_chain(raw, cal, fp64=False, cfa=cfa, demosaic=demosaic)
        err = float(np.abs(got - want).max())
        print(f"{what}: fp32 max abs err {err:.3g}", flush=True)
        assert err <= TOL_FP32, f"{what}: max abs err {err}"
    else:
        want = bench.oracle_chain(raw, cal, fp64=True, cfa=cfa, demosaic=demosaic)
        p = bench.psnr(got, want)
        print(f"{what}: mixed PSNR vs FP64 oracle {p:.2f} dB", flush=True)
        assert p >= PSNR_MIXED_DB, f"{what}: {p} dB"


@pytest.mark.parametrize("precision", ["fp32", "mixed"])
def test_c2_headline_stage_list_on_a_crop_of_the_24mp_frame(gpu_ctx, cal, precision):
    raw = im.synth_raw_u16(bench.W, bench.H, "RGGB", 0)
    crop = np.ascontiguousarray(raw[:CH, :CW])
    got = _gpu_chain(gpu_ctx, crop, "RGGB", cal, precision, "amaze")
    _check(got, crop, "RGGB", cal, precision, "amaze", f"C2 {precision}")


@pytest.mark.parametrize("demosaic", ["markesteijn", "markesteijn3"])
@pytest.mark.parametrize("precision", ["fp32", "mixed"])
def test_c3_xtrans_stage_list_on_a_crop_of_the_26mp_frame(gpu_ctx, cal, precision, demosaic):
    """Demosaic{markesteijn} (the reference's 3x3 mean) and {markesteijn3} (the
    three-pass extension) on X-Trans, then the headline stage list."""
    raw = im.synth_raw_u16(bench.XT_W, bench.XT_H, bench.XTRANS, 3)
    # a crop at the origin keeps the 6x6 pattern's phase; a second one further in
    for y0, x0 in ((0, 0), (1800, 3000)):
        crop = np.ascontiguousarray(raw[y0:y0 + CH, x0:x0 + CW])
        got = _gpu_chain(gpu_ctx, crop, bench.XTRANS, cal, precision, demosaic)
        _check(got, crop, bench.XTRANS, cal, precision, demosaic, f"C3 {precision} {demosaic} crop ({x0},{y0})")


@pytest.mark.parametrize("precision", ["fp32", "mixed"])
def test_c5_burst_crops_through_graph_replays(gpu_ctx, cal, precision):
    """One crop per burst frame size (4000x3000, 6016x4016, 8256x5504 RGGB; 6240x4160
    X-Trans, two of each CFA so every shape's graph replays), per-frame read noise
    drawn as in the bench burst; every frame against the oracle."""
    rng = np.random.default_rng(17)
    frames = []
    for k, (w, h, cfa, _) in enumerate(bench.BURST):
        for j in range(2):
            raw = im.synth_raw_u16(w, h, cfa, 100 + 2 * k + j, read_sigma=float(rng.uniform(1.0, 8.0)))
            y0, x0 = (h // 3) * j, (w // 4) * j
            y0, x0 = y0 - y0 % 6, x0 - x0 % 6  # keeps the Bayer and X-Trans phase
            frames.append((cfa, np.ascontiguousarray(raw[y0:y0 + CH, x0:x0 + CW])))
    for cfa, demosaic in (("RGGB", "amaze"), (bench.XTRANS, "markesteijn")):
        sel = [r for c, r in frames if c == cfa]
        imgs = [im.Image(gpu_ctx).set_calibration_matrix(cal.ravel()).set_raw_u16(r, cfa, bench.BLACK, bench.WHITE)
                for r in sel]
        im.run_burst(imgs, im.parse_config(bench.stages_demosaic(precision, demosaic)), graphs=True)
        for i, (img, r) in enumerate(zip(imgs, sel)):
            _check(img.rgb(), r, cfa, cal, precision, demosaic, f"C5 {precision} {demosaic} frame {i}")
