"""Pins of the markesteijn3 oracle (the three-pass Markesteijn X-Trans
extension; the reference has no such demosaic, so no reference output exists).

What the method's definition fixes: a uniform field and linear grey ramps in
any direction are reproduced exactly (directional linear interpolation of a
linear function, zero colour differences, clamps that do not bind), as is the
green of a linear colour ramp, whose red and blue err only where a colour
difference is extrapolated from one side (bounded); the four
directions are mirror images under transposition (V <-> H, the anti-diagonal
reversed), so transposing a mosaic together with its pattern transposes the
output (bit for bit with one pass; with three passes the candidates' mean is
summed in another order, a few ulps).  Three planted mistakes (a one-column
slip on one diagonal, swapped interpolation weights, a wrong clamp bound) each
fail one of these.  The FP32 form's deviation from FP64 is stated.
"""
import numpy as np
import pytest

import oracle
from imops_amd import synth_raw_u16

XT = "GGRGGB" "GGBGGR" "BRGRBG" "GGBGGR" "GGRGGB" "RBGBRG"
XT_T = "".join(XT[c * 6 + r] for r in range(6) for c in range(6))


def _frame(w, h, seed, pattern=XT):
    return oracle.normalize_raw(synth_raw_u16(w, h, pattern, seed), 600, 16383)


def _failures(plant=0):
    bad = []
    if np.abs(oracle.markesteijn3(np.full((60, 66), 0.5, np.float32), XT, plant=plant) - 0.5).max() != 0:
        bad.append("uniform")
    y, x = np.mgrid[0:120, 0:132].astype(np.float32)
    for name, f in (("x", 0.2 + 0.004 * x), ("y", 0.3 + 0.003 * y), ("xy", 0.1 + 0.002 * x + 0.003 * y),
                    ("anti", 0.7 + 0.002 * x - 0.003 * y)):
        o = oracle.markesteijn3(f.astype(np.float32), XT, plant=plant)
        if np.abs(o - f[..., None])[4:-4, 4:-4].max() > 2.5e-7:
            bad.append("ramp " + name)
    # a linear colour ramp: green is exact; red / blue are exact where the native pixels
    # lie on both sides along the direction and otherwise extrapolate the (linear) colour
    # difference one-sidedly, so the error is bounded (measured 2.3e-4 mean, 1.4e-3 max)
    cols = np.array(["RGB".index(c) for c in XT]).reshape(6, 6)
    ramp = np.stack([0.2 + 0.003 * x + 0.001 * y, 0.3 + 0.001 * x + 0.002 * y, 0.25 + 0.002 * x - 0.001 * y], -1)
    m = np.take_along_axis(ramp, cols[np.arange(120)[:, None] % 6, np.arange(132)[None, :] % 6][..., None], 2)[..., 0]
    o = oracle.markesteijn3(np.ascontiguousarray(m, np.float32), XT, plant=plant)
    d = np.abs(o - ramp)[4:-4, 4:-4]
    if d[..., 1].max() > 2.5e-7 or d.mean() > 2.6e-4 or d.max() > 1.6e-3:
        bad.append("colour ramp")
    p = _frame(150, 150, 1)
    for passes, tol in ((1, 0.0), (3, 4e-7)):
        a = oracle.markesteijn3(p, XT, passes=passes, plant=plant)
        t = oracle.markesteijn3(np.ascontiguousarray(p.T), XT_T, passes=passes, plant=plant).transpose(1, 0, 2)
        if np.abs(a - t).max() > tol:
            bad.append(f"transpose {passes}")
    return bad


def test_markesteijn3_closed_forms_and_transpose_equivariance():
    assert _failures() == []


@pytest.mark.parametrize("plant", [1, 2, 3])
def test_planted_mistakes_are_caught(plant):
    assert _failures(plant), f"planted mistake {plant} passed every pin"


@pytest.mark.parametrize("passes", [1, 3])
def test_markesteijn3_fp32_against_fp64(passes):
    """Error statement: FP32 vs FP64 on a noisy frame; where a homogeneity count or
    a clamp decision flips, the two forms pick different candidates."""
    p = _frame(312, 208, 3)
    a, b = oracle.markesteijn3(p, XT, passes=passes), oracle.markesteijn3(p, XT, passes=passes, fp64=True)
    d = np.abs(a - b)
    assert np.median(d) < 1e-7 and np.mean(d > 1e-5) < 1e-3


def test_markesteijn3_beats_the_3x3_mean_on_resolvable_detail():
    """On a noise-free mosaic of a scene with no detail finer than 6 px the method's
    PSNR exceeds the reference's 3x3 same-colour mean by > 4 dB."""
    h, w = 216, 312
    y, x = np.mgrid[0:h, 0:w].astype(np.float64)
    sc = np.stack([0.4 + 0.2 * np.sin(x / 31 + c) + 0.1 * np.cos(y / 23 + 2 * c) for c in range(3)], -1)
    sc[40:120, 60:200] = (0.7, 0.3, 0.2)
    sc = np.clip(sc + 0.05 * np.sin((x + y) * 2 * np.pi / 9)[..., None], 0, 1).astype(np.float32)
    cols = np.array(["RGB".index(c) for c in XT]).reshape(6, 6)
    mosaic = np.take_along_axis(sc, cols[np.arange(h)[:, None] % 6, np.arange(w)[None, :] % 6][..., None], 2)[..., 0]
    mosaic = np.ascontiguousarray(mosaic, np.float32)
    psnr = lambda o: 10 * np.log10(1 / np.mean((o[8:-8, 8:-8].astype(np.float64) - sc[8:-8, 8:-8]) ** 2))
    assert psnr(oracle.markesteijn3(mosaic, XT)) > psnr(oracle.markesteijn(mosaic, XT)) + 4.0
