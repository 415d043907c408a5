"""Runtime behaviour on the GPU: the device-allocation error path and the
measured-rate probes the benchmark reports.

Allocation failures are injected from the host (imops_debug_fail_allocations):
the library's scratch reservation then fails exactly as a failed hipMalloc
does, so the status code, the message and the recovery of the next call are
exercised without exhausting the device.  The reference's conventions for a
failed call are a non-zero status plus a retrievable last-error message
(extern_pipeline.rs:18-58).
"""
import numpy as np
import pytest

import imops_amd as im
from imops_amd import _lib

pytestmark = pytest.mark.gpu


def test_injected_allocation_failure_reports_and_recovers():
    print("runtime: injected allocation failure", flush=True)
    ctx = im.Context(0)  # a fresh context: nothing reserved yet
    try:
        rgb = im.synth_scene_rgb(64, 48, 3)
        img = im.Image(ctx)
        _lib.check(_lib.lib().imops_debug_fail_allocations(1))
        with pytest.raises(im.ImopsError) as err:
            img.set_rgb(rgb, im.CS_NONE)  # the image's first device buffer
        assert err.value.code == -2  # IMOPS_E_HIP
        assert "hipMalloc" in err.value.message
        # the injection is used up: the same call now succeeds and round-trips the data
        img.set_rgb(rgb, im.CS_NONE)
        np.testing.assert_array_equal(img.rgb(), rgb)
        # a stage's scratch reservation failing mid-chain leaves the context usable
        _lib.check(_lib.lib().imops_debug_fail_allocations(1))
        with pytest.raises(im.ImopsError) as err:
            img.bm3d(0.1)
        assert err.value.code == -2 and "hipMalloc" in err.value.message
        img2 = im.Image(ctx).set_rgb(rgb, im.CS_NONE)
        img2.bm3d(0.1)
        assert np.isfinite(img2.rgb()).all()
    finally:
        _lib.check(_lib.lib().imops_debug_fail_allocations(0))
        ctx.close()


def test_measured_peaks_are_plausible(gpu_ctx):
    print("runtime: measured peaks", flush=True)
    p = gpu_ctx.measure_peaks()
    print(p, flush=True)
    assert 500.0 < p["hbm_copy_GBps"] < 8000.0          # published 8 TB/s
    assert 0.05 < p["valu_fp32_nonfused_instr_per_clk_per_simd"] <= 1.0
    assert 100.0 < p["mfma_f16_TFLOPs"] < 2600.0        # published 2.5 PFLOP/s dense
    assert p["compute_units"] > 0 and p["rated_clock_GHz"] > 0.5


def test_failed_film_preparation_is_not_cached():
    """A film shape whose table upload fails is prepared again on the next call (a cached
    half-prepared entry would hand its kernels missing tables)."""
    print("runtime: film preparation failure", flush=True)
    rgb = (im.synth_scene_rgb(96, 64, 5) * np.float32(40.0)).astype(np.float32)
    ref_ctx = im.Context(0)
    try:
        want = im.Image(ref_ctx).set_rgb(rgb, im.CS_ACESCG).film(stock="Portra400", seed=3).rgb()
    finally:
        ref_ctx.close()
    ctx = im.Context(0)
    try:
        # same frame size first, another stock: the film scratch is then already reserved,
        # so the next allocation is the new stock's prepared tables
        im.Image(ctx).set_rgb(rgb, im.CS_ACESCG).film(stock="ColorNeg200", seed=3)
        img = im.Image(ctx).set_rgb(rgb, im.CS_ACESCG)
        _lib.check(_lib.lib().imops_debug_fail_allocations(1))
        with pytest.raises(im.ImopsError) as err:
            img.film(stock="Portra400", seed=3)
        assert err.value.code == -2 and "hipMalloc" in err.value.message
        got = im.Image(ctx).set_rgb(rgb, im.CS_ACESCG).film(stock="Portra400", seed=3).rgb()
        np.testing.assert_array_equal(got, want)
    finally:
        _lib.check(_lib.lib().imops_debug_fail_allocations(0))
        ctx.close()


def test_bm3d_rejects_frames_wider_than_its_position_format(gpu_ctx):
    """Group tables pack patch positions as x | y << 16: a 65536-pixel-wide frame
    is refused before any kernel runs, and the context stays usable."""
    print("runtime: bm3d width limit", flush=True)
    img = im.Image(gpu_ctx).set_rgb(np.zeros((9, 65536, 3), np.float32), im.CS_NONE)
    with pytest.raises(im.ImopsError) as err:
        img.bm3d(0.1)
    assert err.value.code == _lib.E_INVALID and "65535" in err.value.message
    ok = im.Image(gpu_ctx).set_rgb(im.synth_scene_rgb(64, 48, 3), im.CS_NONE)
    ok.bm3d(0.1)
    assert np.isfinite(ok.rgb()).all()
