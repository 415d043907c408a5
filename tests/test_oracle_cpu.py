"""Pins of the oracle (CPU): the reference's own known-answer tests for the
demosaic/colour path, closed forms and brute force on tiny inputs.

Each check is chosen so that a plausible restatement mistake (a dropped term,
a wrong index, a transposed matrix) fails it.
"""
import numpy as np
import pytest

import oracle

XTRANS = "GGRGGB" "GGBGGR" "BRGRBG" "GGBGGR" "GGRGGB" "RBGBRG"


def _mosaic(scene, pattern="RGGB"):
    h, w, _ = scene.shape
    letters = {"R": 0, "G": 1, "B": 2}
    n = int(round(len(pattern) ** 0.5))
    pw = 2 if len(pattern) == 4 else n
    ph = len(pattern) // pw
    cols = np.array([letters[c] for c in pattern]).reshape(ph, pw)
    idx = cols[np.arange(h)[:, None] % ph, np.arange(w)[None, :] % pw]
    return np.take_along_axis(scene, idx[..., None], 2)[..., 0].astype(np.float32).copy()


@pytest.mark.parametrize("n", [32, 160])
def test_amaze_uniform(n):
    # demosaic.rs:1418-1443 (uniform 0.5 stays within 0.2); the flat field is reproduced exactly
    out = oracle.amaze(np.full((n, n), 0.5, np.float32))
    assert np.all(out == np.float32(0.5))


def test_amaze_linear_ramp_closed_form():
    # a demosaic of a smooth linear ramp reproduces the ramp away from the border
    h = w = 256
    y, x = np.mgrid[0:h, 0:w].astype(np.float32)
    scene = np.stack([0.2 + 0.5 * x / w, 0.3 + 0.4 * y / h, 0.6 - 0.3 * x / w], -1).astype(np.float32)
    out = oracle.amaze(_mosaic(scene))
    assert np.abs(out - scene)[8:-8, 8:-8].max() < 5e-4


def test_amaze_fp32_matches_fp64_oracle():
    from imops_amd import synth_raw_u16
    raw = synth_raw_u16(512, 384, "RGGB", 1)
    plane = oracle.normalize_raw(raw, 600, 16383)
    a32, a64 = oracle.amaze(plane), oracle.amaze(plane, fp64=True)
    d = np.abs(a32 - a64)
    assert np.median(d) < 1e-6
    assert np.mean(d > 1e-5) < 1e-3  # only threshold flips differ


def test_amaze_channel_native_site_passthrough():
    # at a native site the native channel of the output equals the clamped input (output stage 1254-1366)
    from imops_amd import synth_raw_u16
    plane = oracle.normalize_raw(synth_raw_u16(192, 192, "RGGB", 2), 600, 16383)
    out = oracle.amaze(plane)
    g = out[0::2, 1::2, 1]  # G sites of RGGB row 0
    assert np.array_equal(g, np.clip(plane[0::2, 1::2], 0, 1))


def test_markesteijn_brute_force():
    rng = np.random.default_rng(0)
    for pattern, (h, w) in [("RGGB", (7, 9)), (XTRANS, (13, 8))]:
        plane = rng.random((h, w), np.float32)
        out = oracle.markesteijn(plane, pattern)
        pw = 2 if len(pattern) == 4 else 6
        col = lambda r, c: "RGB".index(pattern[(r % (len(pattern) // pw)) * pw + c % pw])
        for r in range(h):
            for c in range(w):
                for ch in range(3):
                    if col(r, c) == ch:
                        want = plane[r, c]
                    else:
                        vals = [plane[y, x] for y in range(max(r - 1, 0), min(r + 1, h - 1) + 1)
                                for x in range(max(c - 1, 0), min(c + 1, w - 1) + 1)
                                if (y, x) != (r, c) and col(y, x) == ch]
                        s = np.float32(0)
                        for v in vals:
                            s = np.float32(s + v)
                        want = np.float32(s / np.float32(len(vals))) if vals else np.float32(0)
                    assert out[r, c, ch] == want


def test_camera_forward_matrix_is_inverse_of_normalised():
    rng = np.random.default_rng(3)
    cal = (np.eye(3) + rng.uniform(-0.2, 0.2, (3, 3))).astype(np.float32)
    fwd = oracle.camera_forward_matrix(cal.ravel()).astype(np.float64)
    d65 = np.array([0.95042, 1.0, 1.08890])
    rgb_d65 = cal.astype(np.float64) @ d65
    norm = cal.astype(np.float64) / rgb_d65[:, None]
    # cst.rs normalises rows by the D65 response, so camera (1,1,1) maps to D65 white
    assert np.allclose(fwd @ norm, np.eye(3), atol=2e-5)
    assert np.allclose(fwd @ np.ones(3), d65, atol=5e-5)


def test_cst_roundtrip_reference_kat():
    # cst.rs:116-128: AcesCg -> Srgb -> AcesCg within 1e-4
    px = np.array([[[0.5, 0.3, 0.8]]], np.float32)
    back = oracle.cst(oracle.cst(px, oracle.CS_ACESCG, oracle.CS_SRGB), oracle.CS_SRGB, oracle.CS_ACESCG)
    assert np.abs(back - px).max() < 1e-4


def test_colour_matrices_white_points():
    # D65 white in XYZ maps to linear sRGB (1,1,1); ACEScg (1,1,1) is D65 white in this matrix set
    white = np.array([[[0.3127 / 0.3290, 1.0, (1 - 0.3127 - 0.3290) / 0.3290]]], np.float32)  # CSS Color 4 D65
    lin = oracle.cst(white, oracle.CS_XYZD65, oracle.CS_LINEAR_SRGB)
    assert np.allclose(lin, 1.0, atol=1e-5)
    aces = oracle.cst(np.ones((1, 1, 3), np.float32), oracle.CS_LINEAR_SRGB, oracle.CS_ACESCG)
    assert np.allclose(aces, 1.0, atol=1e-5)


def test_srgb_oetf_known_values():
    assert oracle.srgb_oetf(0.0) == 0.0
    assert abs(oracle.srgb_oetf(1.0) - 1.0) < 1e-6
    assert abs(oracle.srgb_oetf(0.18) - 0.46135613) < 2e-6   # IEC 61966-2-1
    assert oracle.srgb_oetf(0.001) == np.float32(12.92) * np.float32(0.001)
    assert oracle.srgb_oetf(-0.5) == -oracle.srgb_oetf(0.5)


def test_exp_gamma_wb_closed_forms():
    rgb = np.array([[[0.25, 0.5, 2.0]]], np.float32)
    assert np.array_equal(oracle.exp(rgb, 2.0), rgb * 4)
    assert np.array_equal(oracle.gamma(np.array([[[-1.0, 0.0, 1.0]]], np.float32), 2.2), [[[0, 0, 1]]])
    assert np.allclose(oracle.gamma(rgb, 2.0), np.sqrt(rgb), rtol=1e-6)
    assert np.array_equal(oracle.cfa_coeffs(rgb, [2, 3, 4, 1]), rgb * np.float32([2, 3, 4]))


def test_normalize_raw_crop():
    raw = (np.arange(60, dtype=np.uint16) * 100).reshape(6, 10)
    out = oracle.normalize_raw(raw, 600, 16383, crop=(2, 1, 5, 3))
    want = (raw[1:4, 2:7].astype(np.float32) - np.float32(600)) / np.float32(15783)
    assert np.array_equal(out, want)


# ---------------------------------------------------------------- BM3D pins

def test_bm3d_reference_golden_checksums():
    """bm3d.rs:791-808.  The restatement lands within 0.03 of the recorded Kahan sums
    (~1e-8 relative); evaluating the fixture's sinf/cosf correctly rounded instead of
    with glibc moves the sums by 0.014, and a planted rounding-instead-of-truncation
    in the i64 aggregation moves them by >100, so the tolerance separates libm noise
    from restatement mistakes."""
    img = oracle.bm3d_test_image(2048, 2048)
    out, (refs, grouped) = oracle.bm3d(img, oracle.bm3d_params(0.1))
    assert refs == 2 * 340 * 340
    expected = [2100044.776915211, 2101586.5654235613, 2098591.6187395807]
    for c in range(3):
        assert abs(oracle.refined_sum(out, c) - expected[c]) < 0.05


def test_chroma_bm3d_reference_golden_checksums():
    # bm3d.rs:810-827 (DCT tables, Kaiser window, hard threshold and i64 aggregation without matching)
    img = oracle.bm3d_test_image(2048, 2048)
    out = oracle.chroma_bm3d(img, 0.1)
    expected = [2099732.10061259, 2101692.5347650475, 2098392.755773447]
    for c in range(3):
        assert abs(oracle.refined_sum(out, c) - expected[c]) < 3e-3


def test_bm3d_dct_tables_orthonormal():
    import ctypes as C
    d, i, k = (np.empty(64, np.float32) for _ in range(3))
    oracle.lib().oracle_bm3d_tables(oracle._f(d), oracle._f(i), oracle._f(k))
    D, I = d.reshape(8, 8).astype(np.float64), i.reshape(8, 8).astype(np.float64)
    assert np.allclose(D @ D.T, np.eye(8), atol=1e-6)
    assert np.allclose(I, D.T, atol=1e-6)          # idct[n][k] = dct[k][n]
    K = k.reshape(8, 8)
    assert np.allclose(K, K.T) and np.allclose(K, K[::-1, ::-1])
    assert abs(K[3, 3] - K.max()) < 1e-6 and K.min() > 0


def test_bm3d_fp32_against_fp64_oracle():
    rng = np.random.default_rng(11)
    from imops_amd import synth_scene_rgb
    x = np.clip(synth_scene_rgb(96, 72, 2) + rng.normal(0, 0.03, (72, 96, 3)), 0, 1).astype(np.float32)
    p = oracle.bm3d_params(0.2)
    a, _ = oracle.bm3d(x, p)
    b, _ = oracle.bm3d(x, p, fp64=True)
    mse = float(np.mean((a.astype(np.float64) - b) ** 2))
    assert 10 * np.log10(1.0 / max(mse, 1e-30)) > 60


def test_bm3d_flat_image_is_fixed_point():
    # a flat image has a flat DCT (only DC survives) and is reproduced up to fixed-point rounding
    x = np.full((40, 40, 3), 0.5, np.float32)
    out, _ = oracle.bm3d(x, oracle.bm3d_params(0.3))
    assert np.abs(out - 0.5).max() < 1e-4


def test_preview_demosaic_reference_uniform_and_sites():
    # demosaic.rs:1389-1407: a uniform 16x16 RGGB frame halves to 8x8 of that value
    out = oracle.preview_demosaic(np.full((16, 16), 0.5, np.float32), "RGGB", 2)
    assert out.shape == (8, 8, 3) and np.all(out == np.float32(0.5))
    # the top-left 2x2 cell of each block: RGGB -> R = (0,0), G = the later green site (1,0), B = (1,1)
    a = np.arange(64, dtype=np.float32).reshape(8, 8)
    for f in (2, 4, 8):
        o = oracle.preview_demosaic(a, "RGGB", f)
        for by in range(8 // f):
            for bx in range(8 // f):
                r, c = by * f, bx * f
                assert o[by, bx].tolist() == [a[r, c], a[r + 1, c], a[r + 1, c + 1]]
    # a crop shifted by one column reads the GRBG cell: R = (0,1), G = (1,1), B = (1,0)
    o = oracle.preview_demosaic(a, "RGGB", 2, crop=(1, 0))
    assert o[0, 0].tolist() == [a[0, 1], a[1, 1], a[1, 0]]
