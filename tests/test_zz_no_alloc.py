"""The allocation-free AMaZE entry point (imops_demosaic_amaze_noalloc).

It reads the caller's device CFA plane, writes the caller's device RGB buffer
and keeps its tile scratch in a static device array (64 tile slots); its output
must equal imops_demosaic's (AMaZE) and the oracle's bit for bit.  Two calls in
a row, the second on a smaller frame, with more tiles than slots so that slots
are reused within a call and across calls.
"""
import numpy as np
import pytest
import torch

import imops_amd as im
import oracle
from imops_amd import _lib

pytestmark = pytest.mark.gpu


def _run_noalloc(ctx, plane, pattern, crop=(0, 0)):
    h, w = plane.shape
    d_in = torch.from_numpy(np.ascontiguousarray(plane)).cuda()
    d_out = torch.full((h, w, 3), float("nan"), dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()
    im.demosaic_amaze_noalloc(ctx, d_in.data_ptr(), d_out.data_ptr(), w, h, pattern, crop)
    ctx.synchronize()
    return d_out.cpu().numpy()


def _run_alloc(ctx, plane, pattern, crop=(0, 0)):
    return im.Image(ctx).set_raw_f32(plane, pattern, crop).demosaic("amaze").rgb()


@pytest.mark.parametrize("pattern,crop", [("RGGB", (0, 0)), ("GRBG", (1, 0)), ("BGGR", (1, 1))])
def test_noalloc_amaze_matches_the_allocating_entry_point_twice(gpu_ctx, pattern, crop):
    for k, (w, h, seed) in enumerate([(1056, 864, 3), (480, 288, 4)]):  # 120 then 24 tiles; 64 slots
        print(f"noalloc amaze call {k}: {w}x{h} {pattern} crop {crop}", flush=True)
        raw = im.synth_raw_u16(w, h, pattern, seed)
        plane = oracle.normalize_raw(raw, 600.0, 16383.0)
        got = _run_noalloc(gpu_ctx, plane, pattern, crop)
        want = _run_alloc(gpu_ctx, plane, pattern, crop)
        assert np.array_equal(got.view(np.uint32), want.view(np.uint32)), f"call {k}: differs from imops_demosaic"
        ref = oracle.amaze(plane, pattern, crop_origin=crop)
        assert np.array_equal(got.view(np.uint32), ref.view(np.uint32)), f"call {k}: differs from the oracle"


def test_noalloc_amaze_rejects_what_the_reference_cannot_process(gpu_ctx):
    f = _lib.lib().imops_demosaic_amaze_noalloc
    d = torch.zeros(100 * 200 * 3, dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()
    # width 200: 200 mod 96 = 8 (the reference indexes past its tile buffer there)
    assert f(gpu_ctx.handle, d.data_ptr(), d.data_ptr(), 200, 100, b"RGGB", 0, 0) == _lib.E_UNSUPPORTED
    assert f(gpu_ctx.handle, d.data_ptr(), d.data_ptr(), 192, 96, b"XTRANS?", 0, 0) == _lib.E_INVALID
    assert f(gpu_ctx.handle, None, d.data_ptr(), 192, 96, b"RGGB", 0, 0) == _lib.E_INVALID
