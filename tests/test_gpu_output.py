"""Output encodings on the device (u8 as get_image_rgb_data_c, f16 as save_exr) with
the orientation of apply_orientation, bit-exact against the numpy oracle.  Inputs
cover negatives, values above 1, NaN and infinities, the f16 subnormal range, the
largest finite f16 and overflow, and values on the u8 rounding boundaries."""
import numpy as np
import pytest

import imops_amd as im
import oracle

pytestmark = pytest.mark.gpu


def _values(w, h, seed):
    rng = np.random.default_rng(seed)
    x = rng.uniform(-0.2, 1.3, (h, w, 3)).astype(np.float32)
    special = np.array([np.nan, np.inf, -np.inf, 0.0, -0.0, 1.0, 254.5 / 255, 1 / 255, 6e-8, 3e-5, 6.1e-5,
                        65504.0, 65520.0, 1e6, 0.5 + 1 / 4096, 2.0 ** -24, 2.0 ** -25], np.float32)
    flat = x.reshape(-1)
    k = min(special.size, flat.size)
    flat[:k] = special[:k]
    return x


@pytest.mark.parametrize("orientation", [0, 1, 2, 3])
@pytest.mark.parametrize("w,h", [(37, 23), (1, 5), (64, 64)])
def test_encodings_and_orientation(gpu_ctx, w, h, orientation):
    x = _values(w, h, w * 10 + h + orientation)
    img = im.Image(gpu_ctx).set_rgb(x, im.CS_SRGB)
    u8 = img.rgb_u8(orientation)
    assert np.array_equal(u8, oracle.orient(oracle.encode_u8(x), orientation))
    f16 = img.rgb_f16(orientation)
    want = oracle.orient(oracle.encode_f16(x), orientation)
    assert np.array_equal(f16.view(np.uint16), want.view(np.uint16))


def test_encoding_rejects_bad_orientation(gpu_ctx):
    img = im.Image(gpu_ctx).set_rgb(np.zeros((2, 2, 3), np.float32), im.CS_SRGB)
    with pytest.raises(im.ImopsError, match="orientation"):
        img.rgb_u8(4)
