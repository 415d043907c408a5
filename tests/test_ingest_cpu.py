"""Host-side raw ingest of the C ABI (no kernel launches): the calibration
matrix parse_raw_image picks or interpolates between the illuminant-A and
D65 colour matrices (pichromatic-pipeline/src/extern_pipeline.rs:254-270,
interpolate_matrices 457-495).  Pinned by its special cases, which the
interpolation fixes exactly: the white balance's B/R ratio at illuminant A
gives matrix A, at D65 (or an unusable white balance) matrix D65; the
geometric mean of the two ratios gives their midpoint; the order of the
fallbacks."""
import numpy as np
import pytest

import imops_amd as im

F = np.float32
A = np.array([0.6, 0.1, -0.05, -0.3, 1.2, 0.1, -0.02, 0.15, 0.55], F)
D = np.array([0.8, -0.2, -0.1, -0.4, 1.3, 0.05, -0.1, 0.3, 0.7], F)


def _ratio(m, white):
    # s = M . white in the reference's f32 order; ratio = s[0] / s[2]
    s = [F(F(m[i * 3] * F(white[0])) + F(m[i * 3 + 1] * F(white[1]))) + F(m[i * 3 + 2] * F(white[2])) for i in range(3)]
    return F(s[0] / s[2])


R1 = _ratio(A, (1.09850, 1.0, 0.35585))
R2 = _ratio(D, (0.95047, 1.0, 1.08883))


def test_fallback_order():
    I = np.eye(3, dtype=F).ravel()
    assert np.array_equal(im.dng_calibration_matrix(), I)
    assert np.array_equal(im.dng_calibration_matrix(color_matrix_a=A), A)
    assert np.array_equal(im.dng_calibration_matrix(color_matrix_d65=D), D)
    assert np.array_equal(im.dng_calibration_matrix(color_matrix_other=A), A)
    assert np.array_equal(im.dng_calibration_matrix(color_matrix_d65=D, color_matrix_other=A), D)


@pytest.mark.parametrize("wb", [None, (0, 1, 2, 1), (2, 1, 0, 1), (float("nan"), 1, 1.5, 1)])
def test_unusable_white_balance_gives_the_d65_matrix(wb):
    got = im.dng_calibration_matrix(A, D, wb_coeffs=wb)
    assert np.array_equal(got, D)  # ratio_s = ratio2: weight exactly 0


def test_white_balance_at_illuminant_a_gives_matrix_a():
    got = im.dng_calibration_matrix(A, D, wb_coeffs=(1.0, 1.0, float(R1), 1.0))
    assert np.array_equal(got, A)  # weight exactly 1


def test_geometric_mean_ratio_gives_the_midpoint_and_weights_clamp():
    mid = float(np.sqrt(np.float64(R1) * np.float64(R2)))
    got = im.dng_calibration_matrix(A, D, wb_coeffs=(1.0, 1.0, mid, 1.0))
    np.testing.assert_allclose(got, 0.5 * A.astype(np.float64) + 0.5 * D.astype(np.float64), atol=2e-6)
    beyond = float(R1) * 4 if R1 > R2 else float(R1) / 4  # past illuminant A: clamped to matrix A
    assert np.array_equal(im.dng_calibration_matrix(A, D, wb_coeffs=(1.0, 1.0, beyond, 1.0)), A)
