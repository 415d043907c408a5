"""The oracle's Haar group transform (the BM3D group_transform = "haar" option;
no reference counterpart: bm3d.rs:595-614 fixes the Walsh-Hadamard transform).

Pins: the forward transform of every unit vector equals the textbook
orthonormal Haar basis built by the Kronecker recursion
H_2m = [H_m (x) (1, 1); I_m (x) (1, -1)] / sqrt(2), rows mapped from the
in-place lifting order (position k > 0 with lowest set bit h holds the detail
of block (k - h) / 2h at support 2h, which is textbook row n / 2h + (k - h) / 2h);
inverse o forward is the identity; for groups of two it equals the
Walsh-Hadamard transform bit for bit; and through the whole BM3D a flat image
is a fixed point, groups of two give the Walsh-Hadamard output, and the FP32
form stays > 60 dB from the FP64 one.  A wrong scale, sign, pairing stride or
stage order changes some row of the basis and fails the first test.
"""
import numpy as np
import pytest

import imops_amd as im
import oracle


def textbook_haar(n):
    h = np.array([[1.0]])
    m = 1
    while m < n:
        h = np.vstack([np.kron(h, [1.0, 1.0]), np.kron(np.eye(m), [1.0, -1.0])]) / np.sqrt(2.0)
        m *= 2
    return h


def lifting_row(k, n):
    if k == 0:
        return 0
    h = k & -k
    return n // (2 * h) + (k - h) // (2 * h)


@pytest.mark.parametrize("n", [1, 2, 4, 8, 16, 32])
def test_haar_forward_is_the_textbook_basis(n):
    f = np.stack([oracle.group_transform_1d(e, "haar") for e in np.eye(n)], axis=1)  # column i = F e_i
    perm = [lifting_row(k, n) for k in range(n)]
    assert sorted(perm) == list(range(n))
    np.testing.assert_allclose(f, textbook_haar(n)[perm], atol=1e-12)
    np.testing.assert_allclose(f @ f.T, np.eye(n), atol=1e-12)


@pytest.mark.parametrize("n", [2, 4, 8, 16])
def test_haar_inverse_round_trip(n):
    rng = np.random.default_rng(n)
    x = rng.normal(size=n) * 100
    for fp64, tol in ((True, 1e-12), (False, 1e-3)):
        y = oracle.group_transform_1d(oracle.group_transform_1d(x, "haar", fp64=fp64), "haar", inverse=True, fp64=fp64)
        np.testing.assert_allclose(y, x, atol=tol)


def test_haar_equals_hadamard_for_two():
    rng = np.random.default_rng(3)
    for _ in range(50):
        x = rng.normal(size=2).astype(np.float32).astype(np.float64) * 50
        for fp64 in (False, True):
            a = oracle.group_transform_1d(x, "haar", fp64=fp64)
            b = oracle.group_transform_1d(x, "hadamard", fp64=fp64)
            assert np.array_equal(a, b)
            assert np.array_equal(oracle.group_transform_1d(a, "haar", inverse=True, fp64=fp64),
                                  oracle.group_transform_1d(b, "hadamard", inverse=True, fp64=fp64))


def _scene(w, h, seed, sigma=0.03):
    rng = np.random.default_rng(seed)
    return np.clip(im.synth_scene_rgb(w, h, seed) + rng.normal(0, sigma, (h, w, 3)), 0, 1).astype(np.float32)


def test_bm3d_haar_flat_fixed_point_and_fp64():
    x = np.full((40, 40, 3), 0.5, np.float32)
    out, _ = oracle.bm3d(x, oracle.bm3d_params(0.3, group_transform="haar"))
    assert np.abs(out - 0.5).max() < 1e-4
    rgb = _scene(96, 72, 2)
    p = oracle.bm3d_params(0.2, group_transform="haar")
    a, _ = oracle.bm3d(rgb, p)
    b, _ = oracle.bm3d(rgb, p, fp64=True)
    mse = float(np.mean((a.astype(np.float64) - b) ** 2))
    assert 10 * np.log10(1.0 / max(mse, 1e-30)) > 60
    c, _ = oracle.bm3d(rgb, oracle.bm3d_params(0.2))
    assert not np.array_equal(a, c)  # the option changes the result for groups of four or more


def test_bm3d_haar_groups_of_two_equal_hadamard():
    rgb = _scene(64, 56, 4)
    a, _ = oracle.bm3d(rgb, oracle.bm3d_params(0.3, group_transform="haar", max_group=2))
    b, _ = oracle.bm3d(rgb, oracle.bm3d_params(0.3, max_group=2))
    assert np.array_equal(a, b)


def test_group_transform_stage_key():
    cfg = im.parse_config('[[pipeline_modules]]\nname = "BM3D"\nintensity = 0.1\ngroup_transform = "haar"\n')
    assert cfg.pipeline_modules == ["BM3D"]
    with pytest.raises(im.ImopsError):
        im.parse_config('[[pipeline_modules]]\nname = "BM3D"\ngroup_transform = "daub4"\n')
    p = im.bm3d_params(0.1, group_transform="haar")
    assert p.group_transform == 1 and im.bm3d_params(0.1).group_transform == 0
