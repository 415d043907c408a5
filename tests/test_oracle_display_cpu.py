"""CPU pins of the oracle's display-tail stages (contrast.rs, highlight_reconstruction.rs,
lch.rs, tone_map.rs, vignette.rs, BaselineExposureCompensation in modules.rs:335-371).

Pins: the reference's own unit tests (highlight_reconstruction.rs:17-33, lch.rs:23-36,
tone_map.rs:172-183), the published Oklab values of the sRGB primaries and white
(CSS Color 4 / Ottosson), the defining properties of the ACES reference gamut
compression curve (identity below the threshold, the limit mapped onto the gamut
boundary, monotone), and closed forms for contrast, exposure gain and the vignette
polynomial.  The colour crate's own Oklch code is absent, so its exact rounding is
unpinned.
"""
import struct

import numpy as np
import pytest

import oracle


def test_highlight_reconstruction_reference_case():
    out = oracle.highlight(np.array([[[0.6, 0.4, 0.8]]], np.float32), 1.0)
    assert np.abs(out[0, 0] - [0.6, 0.52, 0.8]).max() < 1e-6


def test_lch_reference_roundtrip():
    px = np.array([[[0.4, 0.7, 0.2]]], np.float32)
    out = oracle.lch(px, oracle.CS_SRGB, 1.0, 1.0, 1.0)
    assert np.abs(out - px).max() < 1e-4


@pytest.mark.parametrize("lin,lab", [((1, 1, 1), (1.0, 0.0, 0.0)), ((1, 0, 0), (0.62796, 0.22486, 0.12585)),
                                     ((0, 1, 0), (0.86644, -0.23389, 0.17950)),
                                     ((0, 0, 1), (0.45201, -0.03246, -0.31153))])
def test_oklab_published_values(lin, lab):
    L, C, h = oracle.oklch(np.array(lin, np.float32))
    assert abs(L - lab[0]) < 1e-4
    assert abs(C - np.hypot(lab[1], lab[2])) < 1e-4
    if C > 1e-3:
        want = np.degrees(np.arctan2(lab[2], lab[1])) % 360
        assert abs(h - want) < 0.02


def test_sigmoid_keeps_middle_grey():
    out = oracle.sigmoid(np.full((1, 1, 3), 0.18, np.float32))
    assert np.abs(out - 0.18).max() < 1e-4


def test_gamut_compression_curve():
    thr, lim = [0.815, 0.803, 0.88], [1.147, 1.264, 1.312]
    for c in range(3):
        below = np.linspace(0, thr[c] - 1e-3, 7)
        assert all(oracle.lib().oracle_rgc_curve(float(d), c) == np.float32(d) for d in below)
        assert abs(oracle.lib().oracle_rgc_curve(lim[c], c) - 1.0) < 1e-5
        ys = [oracle.lib().oracle_rgc_curve(float(d), c) for d in np.linspace(thr[c], 3.0, 50)]
        assert all(b > a for a, b in zip(ys, ys[1:]))
    # in-gamut colours pass unchanged
    px = np.array([0.5, 0.3, 0.2], np.float32)
    assert np.abs(oracle.rgc(px) - px).max() < 1e-6


def test_contrast_and_exposure_closed_forms():
    x = np.random.default_rng(0).uniform(0.01, 2.0, (4, 5, 3)).astype(np.float32)
    assert np.abs(oracle.contrast(x, 1.0) - x).max() < 1e-6
    want = 0.185 * (x.astype(np.float64) / 0.185) ** 2
    assert np.abs(oracle.contrast(x, 2.0) / want - 1).max() < 1e-6
    g = 12.5 * 8.0 ** 2 / (1 / 250 * 400)
    assert np.abs(oracle.baseline(x, 0.0, 1 / 250, 8.0, 400.0) / (x * g) - 1).max() < 1e-6
    assert np.abs(oracle.baseline(x, 1.0, 0.0, 0.0, 0.0) - 2 * x).max() < 1e-6


def vignette_opcodes(k, cx, cy, extra_first=False):
    body = struct.pack(">7d", *k, cx, cy)
    ops = []
    if extra_first:
        ops.append(struct.pack(">IIII", 1, 1, 0, 8) + b"\0" * 8)  # another opcode, skipped
    ops.append(struct.pack(">IIII", 3, 1, 0, len(body)) + body)
    return struct.pack(">I", len(ops)) + b"".join(ops)


def test_vignette_polynomial_closed_form():
    w, h = 41, 29
    x = np.ones((h, w, 3), np.float32)
    k = (-0.3, 0.1, 0.0, 0.0, 0.05)
    out = oracle.vignette(x, vignette_opcodes(k, 0.5, 0.4, extra_first=True), 0.8)
    yy, xx = np.mgrid[0:h, 0:w]
    d = np.hypot(xx / (w - 1) - 0.5, yy / (h - 1) - 0.4)
    dmax = max(np.hypot(a - 0.5, b - 0.4) for a in (0, 1) for b in (0, 1))
    r = d / dmax
    gain = np.maximum(1 + 0.8 * sum(kk * r ** (2 * (i + 1)) for i, kk in enumerate(k)), 0)
    assert np.abs(out[..., 0] - gain).max() < 1e-5
    # a truncated list does nothing
    assert np.array_equal(oracle.vignette(x, vignette_opcodes(k, 0.5, 0.4)[:-3], 0.8), x)
