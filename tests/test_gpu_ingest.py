"""Raw ingest on the device: a bound u16 frame read and normalised inside
AMaZE's tile fill (crop_and_normalize, demosaic.rs:61-91, with the crop
origin's CFA shift, demosaic.rs:93-98) gives the bits of the copying path and
of the oracle; other consumers convert it first; crop_bayer_center
(extern_pipeline.rs:113-179) against a numpy brute force."""
import numpy as np
import pytest
import torch

import imops_amd as im
import oracle

pytestmark = pytest.mark.gpu

XT = "GGRGGB" "GGBGGR" "BRGRBG" "GGBGGR" "GGRGGB" "RBGBRG"


def _dev(raw):
    t = torch.from_numpy(raw.view(np.int16)).cuda()
    torch.cuda.synchronize()
    return t


@pytest.mark.parametrize("w,h,crop", [(480, 288, None), (500, 320, (3, 5, 384, 288)), (1056, 864, (1, 0, 960, 768))])
def test_bound_u16_amaze_equals_the_copying_path_and_the_oracle(gpu_ctx, w, h, crop):
    raw = im.synth_raw_u16(w, h, "RGGB", 4)
    d = _dev(raw)
    a = im.Image(gpu_ctx).bind_raw_u16(d, "RGGB", 600.0, 16383.0, crop).demosaic("amaze").rgb()
    b = im.Image(gpu_ctx).set_raw_u16(raw, "RGGB", 600.0, 16383.0, crop).demosaic("amaze").rgb()
    assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
    cx, cy, cw, ch = crop if crop else (0, 0, w, h)
    plane = oracle.normalize_raw(raw, 600.0, 16383.0, (cx, cy, cw, ch))
    want = oracle.amaze(plane, "RGGB", crop_origin=(cx, cy))
    assert np.array_equal(a.view(np.uint32), want.view(np.uint32))


@pytest.mark.parametrize("algo", ["markesteijn", "markesteijn3", "fast"])
def test_bound_u16_other_demosaics_convert_first(gpu_ctx, algo):
    raw = im.synth_raw_u16(312, 208, XT if algo != "fast" else "RGGB", 5)
    cfa = XT if algo != "fast" else "RGGB"
    a = im.Image(gpu_ctx).bind_raw_u16(_dev(raw), cfa, 600.0, 16383.0).demosaic(algo).rgb()
    b = im.Image(gpu_ctx).set_raw_u16(raw, cfa, 600.0, 16383.0).demosaic(algo).rgb()
    assert np.array_equal(a.view(np.uint32), b.view(np.uint32))


def test_bound_frames_in_a_graph_burst(gpu_ctx):
    cfg = im.parse_config('[[pipeline_modules]]\nname = "Demosaic"\nalgorithm = "amaze"\n'
                          '[[pipeline_modules]]\nname = "CST"\ntarget_color_space = "Srgb"\n')
    cal = np.array([[1.1, -0.1, 0.05], [-0.08, 0.95, 0.1], [0.02, -0.12, 1.2]], np.float32).ravel()
    raws = [im.synth_raw_u16(288, 192, "RGGB", 40 + k) for k in range(3)]
    devs = [_dev(r) for r in raws]
    frames = [im.Image(gpu_ctx).set_calibration_matrix(cal).bind_raw_u16(d, "RGGB", 600.0, 16383.0) for d in devs]
    im.run_burst(frames, cfg, graphs=True)
    for r, f in zip(raws, frames):
        ref = im.Image(gpu_ctx).set_calibration_matrix(cal).set_raw_u16(r, "RGGB", 600.0, 16383.0)
        im.run_pixel_pipeline(ref, cfg)
        assert np.array_equal(f.rgb().view(np.uint32), ref.rgb().view(np.uint32))


@pytest.mark.parametrize("w,h,factor", [(200, 150, 2), (201, 151, 3), (96, 64, 4), (64, 48, 1)])
def test_crop_bayer_center(gpu_ctx, w, h, factor):
    plane = np.random.default_rng(w).random((h, w), np.float32)
    img = im.Image(gpu_ctx).set_raw_f32(plane, "RGGB").crop_bayer_center(factor)
    if factor <= 1:
        assert (img.width, img.height) == (w, h)
        return
    nw, nh = (w // factor) & ~1, (h // factor) & ~1
    assert (img.width, img.height) == (nw, nh)
    ys = (np.arange(nh) // 2) * factor * 2 + np.arange(nh) % 2
    xs = (np.arange(nw) // 2) * factor * 2 + np.arange(nw) % 2
    want = plane[ys[:, None], xs[None, :]]
    got = img.demosaic("markesteijn").rgb()
    ref = oracle.markesteijn(np.ascontiguousarray(want), "RGGB")
    assert np.array_equal(got.view(np.uint32), ref.view(np.uint32))
