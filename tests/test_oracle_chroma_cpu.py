"""CPU pins of the oracle's luma-guided chroma denoise (chroma_denoise.rs:14-152).

The reference holds no golden vectors for this stage, so the restatement is
pinned by what the mathematics fixes: the running-sum box filter against a
brute-force f64 mean over the in-bounds window (an off-by-one in the window or
the normalisation moves values by ~1e-1, the f32 running-sum drift is ~1e-6),
the guided filter against an f64 restatement built on that brute-force mean,
and two invariants: a grey image has no chroma to smooth, and a flat image
stays flat.
"""
import numpy as np
import pytest

import oracle

WR, WG, WB = 0.2722287168, 0.6740817658, 0.0536895174


def _box64(a, r):
    h, w = a.shape
    out = np.empty((h, w))
    for y in range(h):
        for x in range(w):
            out[y, x] = a[max(0, y - r):min(h, y + r + 1), max(0, x - r):min(w, x + r + 1)].mean()
    return out


def _guided64(g, p, r, eps):
    mg, mp = _box64(g, r), _box64(p, r)
    var = _box64(g * g, r) - mg * mg
    cov = _box64(g * p, r) - mg * mp
    a = cov / (var + eps)
    b = mp - a * mg
    return _box64(a, r) * g + _box64(b, r)


def _denoise64(rgb, r, eps):
    x = rgb.astype(np.float64)
    y = WR * x[..., 0] + WG * x[..., 1] + WB * x[..., 2]
    qr, qb = _guided64(y, x[..., 0] - y, r, eps), _guided64(y, x[..., 2] - y, r, eps)
    return np.stack([y + qr, y - WR / WG * qr - WB / WG * qb, y + qb], -1)


@pytest.mark.parametrize("w,h,r", [(37, 23, 0), (37, 23, 1), (37, 23, 4), (19, 31, 7), (12, 9, 50), (1, 17, 3)])
def test_box_filter_is_the_in_bounds_window_mean(w, h, r):
    a = np.random.default_rng(w * h + r).uniform(0, 1, (h, w)).astype(np.float32)
    got = oracle.box_filter(a, r)
    assert np.abs(got - _box64(a.astype(np.float64), r)).max() < 2e-6


@pytest.mark.parametrize("r,eps", [(2, 0.01), (4, 0.01), (3, 0.1)])
def test_guided_denoise_matches_f64_restatement(r, eps):
    rng = np.random.default_rng(r)
    rgb = np.clip(rng.uniform(0.1, 0.9, (1, 1, 3)) + rng.normal(0, 0.08, (29, 33, 3)), 0, 1).astype(np.float32)
    got = oracle.luma_chroma_denoise(rgb, r, eps)
    want = _denoise64(rgb, r, eps)
    assert np.abs(got - want).max() < 2e-5
    # the stage does smooth chroma: the output's chroma varies less than the input's
    chroma = lambda x: x[..., 0] - (WR * x[..., 0] + WG * x[..., 1] + WB * x[..., 2])
    assert chroma(got).std() < 0.8 * chroma(rgb).std()


def test_grey_and_flat_images_are_kept():
    v = np.random.default_rng(1).uniform(0, 1, (20, 24, 1)).astype(np.float32)
    grey = np.repeat(v, 3, axis=2)
    assert np.abs(oracle.luma_chroma_denoise(grey, 4, 0.01) - grey).max() < 1e-6
    flat = np.broadcast_to(np.array([0.7, 0.2, 0.4], np.float32), (16, 21, 3)).copy()
    assert np.abs(oracle.luma_chroma_denoise(flat, 3, 0.01) - flat).max() < 1e-6
