"""GPU parity of the three-pass Markesteijn demosaic (markesteijn3) against the
FP32 oracle, bit for bit (the kernels follow the oracle's operation order), on
X-Trans and Bayer mosaics, odd sizes and crop-shifted patterns, one and three
passes; through the stage list's Demosaic{algorithm = "markesteijn3"}."""
import numpy as np
import pytest

import imops_amd as im
import oracle

pytestmark = pytest.mark.gpu

XT = "GGRGGB" "GGBGGR" "BRGRBG" "GGBGGR" "GGRGGB" "RBGBRG"


@pytest.mark.parametrize("pattern,w,h,crop,seed", [(XT, 312, 208, (0, 0), 1), (XT, 253, 171, (2, 5), 2),
                                                    (XT, 96, 64, (5, 1), 3), ("RGGB", 200, 150, (0, 0), 4),
                                                    ("GRBG", 131, 77, (1, 0), 5)])
@pytest.mark.parametrize("passes", [1, 3])
def test_markesteijn3_bit_exact(gpu_ctx, pattern, w, h, crop, seed, passes):
    print(f"markesteijn3 {w}x{h} {pattern[:4]} crop {crop} passes {passes}", flush=True)
    plane = oracle.normalize_raw(im.synth_raw_u16(w, h, pattern, seed), 600, 16383)
    algo = "markesteijn3" if passes == 3 else "markesteijn3_1pass"
    got = im.Image(gpu_ctx).set_raw_f32(plane, pattern, crop).demosaic(algo).rgb()
    want = oracle.markesteijn3(plane, pattern, crop_origin=crop, passes=passes)
    bad = np.count_nonzero(got.view(np.uint32) != want.view(np.uint32))
    assert bad == 0, f"{bad} values differ, max {np.abs(got - want).max()}"


def test_markesteijn3_through_the_stage_list(gpu_ctx):
    raw = im.synth_raw_u16(300, 200, XT, 7)
    img = im.Image(gpu_ctx).set_raw_u16(raw, XT, 600.0, 16383.0)
    im.run_pixel_pipeline(img, im.parse_config('[[pipeline_modules]]\nname = "Demosaic"\nalgorithm = "markesteijn3"\n'))
    want = oracle.markesteijn3(oracle.normalize_raw(raw, 600, 16383), XT)
    assert np.array_equal(img.rgb().view(np.uint32), want.view(np.uint32))
    img1 = im.Image(gpu_ctx).set_raw_u16(raw, XT, 600.0, 16383.0)
    im.run_pixel_pipeline(img1, im.parse_config(
        '[[pipeline_modules]]\nname = "Demosaic"\nalgorithm = "markesteijn3"\npasses = 1\n'))
    want1 = oracle.markesteijn3(oracle.normalize_raw(raw, 600, 16383), XT, passes=1)
    assert np.array_equal(img1.rgb().view(np.uint32), want1.view(np.uint32))
    with pytest.raises(im.ImopsError):
        im.parse_config('[[pipeline_modules]]\nname = "Demosaic"\nalgorithm = "markesteijn3"\npasses = 2\n')
